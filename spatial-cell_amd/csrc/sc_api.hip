// C ABI of the gfx950 pipeline (include/spatial_cell.h).
//
// One step enqueues the kernels of spatial_cell_sim's loop body
// (main.cu:347-649) on the handle's stream with every count kept on the
// device, so a step never waits on the host (the reference synchronises ~98
// times per step and reads nActive back three times, main.cu:369,399,616).
// By default each step is replayed from a captured hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "sc_batch.h"
#include "sc_device.h"
#include "sc_setup.h"

using namespace scd;

static thread_local std::string g_create_error;

namespace scd {
// message of a failed call without a handle (sc_bin), read by sc_last_error(NULL)
void set_handleless_error(const std::string& e) { g_create_error = e; }
}


struct GraphEntry {
  hipGraphExec_t exec;
  uint64_t end_sig;
};

struct sc_sim {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream of a step: launches that do not depend on the main
  // stream's run beside them (fork/join through the events below)
  hipStream_t side = nullptr;
  std::vector<hipEvent_t> fork_ev;
  // 0: one chain of launches (default: measured fastest, the graph's
  // parallel branches cost more than they overlap); 1: population branches;
  // 2: the same as 1 (bonded particles now share the relax pass kernels)
  int concurrent = 0;
  // while a whole step is queued with two relax kernels (relax_split): the
  // fork that runs each pass's general kernel beside its collision-only one
  Fork* relax_fork = nullptr;
  // asynchronous frames: device-to-host copies of the reduced frame run on
  // their own stream; frame_ev marks the last one's end
  hipStream_t copy = nullptr;
  hipEvent_t frame_ev = nullptr;
  // fused relaxation of small components (option "fuse"); `fusing` is set
  // while a whole step is queued (the stage-by-stage path never fuses)
  Fuse fuse{};
  bool fuse_on = false;
  uint32_t* pool_ctrs = nullptr;   // pool counters of the four pair lists, kPoolShards each
  // batch: scratch of instance_slots (tile counts, slot list, total) and
  // the staging of download_slots
  int* islots = nullptr;
  size_t islots_cap = 0;
  void* gather_buf = nullptr;
  size_t gather_cap = 0;
  hipEvent_t retire_ev = nullptr;   // server: a retiring instance's gather is done
  // batch: staging of upload_slots (device, pinned host)
  void* up_dev = nullptr;
  void* up_host = nullptr;
  size_t up_cap = 0;
  bool fusing = false;
  bool site_rows = false;   // some interaction row bonds by site contact (bond_probability > 0)
  scs::HostSetup host;
  GridP g;
  Phys ph;
  Tables t;
  Pop pop[2];
  Counters* ctr = nullptr;
  std::vector<void*> allocs;
  std::string err;
  bool use_graph = true;
  std::map<uint64_t, GraphEntry> graphs;
  void* d_frame = nullptr;
  size_t frame_bytes = 0;
  // batch: the instances' setups (host = instance 0's, for the shared
  // constants), the per-instance constant table and counters on the device
  bool batch = false;
  std::vector<scs::HostSetup> inst;
  std::vector<GridP> table;
  GridP* d_table = nullptr;
  InstCtr* d_ictr = nullptr;
  // serving: instance slots in use (live or evicted and not yet dropped)
  bool server = false;
  std::vector<char> live;
  // list pool demand of the last step (pinned copy of the shard counters)
  uint32_t* h_demand = nullptr;
  hipEvent_t demand_ev = nullptr;
  bool demand_pending = false;
};

#define SC_TRY(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      sim->err = std::string(#expr) + ": " + hipGetErrorString(e_);                    \
      return e_ == hipErrorOutOfMemory ? SC_ERR_OOM : SC_ERR_DEVICE;                   \
    }                                                                                  \
  } while (0)

static sc_status fail(sc_sim* sim, sc_status s, const std::string& m) {
  if (sim) sim->err = m;
  return s;
}

// Host copies and fills run on the handle's own non-blocking stream and are
// waited for.  The legacy default stream is never used: it orders against
// the blocking streams of the process, and HIP refuses it while another
// thread captures a graph (a serving loop beside the caller's own
// simulations, for one).
static hipError_t mem_copy(sc_sim* sim, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, sim->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(sim->stream);
}
static hipError_t mem_fill(sc_sim* sim, void* dst, int value, size_t bytes) {
  hipError_t e = hipMemsetAsync(dst, value, bytes, sim->stream);
  return e != hipSuccess ? e : hipStreamSynchronize(sim->stream);
}

static sc_status dalloc(sc_sim* sim, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) { *p = nullptr; return fail(sim, SC_ERR_OOM, std::string("hipMalloc failed: ") + hipGetErrorString(e)); }
  sim->allocs.push_back(*p);
  SC_TRY(mem_fill(sim, *p, 0, bytes));
  return SC_OK;
}

static size_t field_bytes(Field f, size_t cap, bool metabolic) {
  switch (f) {
    case F_BONDS: return metabolic ? 0 : cap * SC_MAX_BONDS * sizeof(int4);
    case F_MET0: case F_MET1: return metabolic ? cap * sizeof(float4) : 0;
    case F_MET2: return metabolic ? cap * sizeof(float2) : 0;
    case F_HDR: case F_ST: return cap * sizeof(int4);
    case F_INST: return cap * sizeof(int);
    default: return cap * sizeof(float4);
  }
}

// Union-find and packing arrays of the fused relaxation over both arrays' capacities.
static sc_status alloc_fuse(sc_sim* sim) {
  Fuse& f = sim->fuse;
  f.pcap = sim->pop[0].cap;
  f.nodes = sim->pop[0].cap + sim->pop[1].cap;
  f.tiles = (f.nodes + kFuseTile - 1) / kFuseTile;
  if (f.tiles < 1) f.tiles = 1;
  const size_t n = (size_t)(f.nodes > 0 ? f.nodes : 1);
  sc_status s;
  int** per_node[] = {&f.parent, &f.size, &f.flags, &f.offset, &f.fill, &f.loc};
  for (int** a : per_node)
    if ((s = dalloc(sim, (void**)a, n * sizeof(int))) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&f.members, (size_t)f.tiles * kFuseCap * sizeof(int))) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&f.count, (size_t)f.tiles * kFuseWaves * sizeof(int))) != SC_OK) return s;
  return dalloc(sim, (void**)&f.any_walk, sizeof(int));
}

static sc_status alloc_pop(sc_sim* sim, Pop& p, int cap, bool metabolic, size_t cells) {
  std::memset(&p, 0, sizeof(p));
  p.cap = cap;
  p.launch = cap > 0 ? cap : 1;
  p.metabolic = metabolic;
  size_t c = (size_t)(cap > 0 ? cap : 1);
  for (int f = 0; f < F_COUNT; f++) {
    size_t b = field_bytes((Field)f, c, metabolic);
    for (int k = 0; k < 2; k++) {
      if (b == 0) { p.buf[f][k] = nullptr; continue; }
      sc_status s = dalloc(sim, &p.buf[f][k], b);
      if (s != SC_OK) return s;
    }
    p.cur[f] = 0;
  }
  sc_status s;
  if ((s = dalloc(sim, (void**)&p.keys_in, c * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.keys_tmp, c * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.keys_sorted, c * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.vals_tmp, 2 * c * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.vals_sorted, c * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.cs, (cells + 1) * sizeof(uint32_t))) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.gaps, (cells / 64 + 2) * sizeof(uint4))) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.gap_ctr, sizeof(uint32_t))) != SC_OK) return s;
  p.max_tiles = (int)((c + kBlock * kSortItemsMin - 1) / (kBlock * kSortItemsMin));   // the smallest tiles
  if ((s = dalloc(sim, (void**)&p.hist, (size_t)kMaxPasses * kRadix * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.status, (size_t)kMaxPasses * p.max_tiles * kRadix * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.tile_ctr, kMaxPasses * 4)) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.flw, c * sizeof(uint32_t))) != SC_OK) return s;
  if (!metabolic && (s = dalloc(sim, (void**)&p.site, c * 4 * sizeof(int2))) != SC_OK) return s;
  if (!metabolic && (s = dalloc(sim, (void**)&p.bnb, c * kBondNbrs * sizeof(uint2))) != SC_OK) return s;
  if (!metabolic && (s = dalloc(sim, (void**)&p.rbp, c * kBondNbrs * sizeof(uint2))) != SC_OK) return s;
  if (!metabolic && (s = dalloc(sim, (void**)&p.bterm, c * SC_MAX_BONDS * sizeof(float4))) != SC_OK) return s;
  if (!metabolic && (s = dalloc(sim, (void**)&p.bang, c * sizeof(float4))) != SC_OK) return s;
  for (int k = 0; k < 2 && !metabolic; k++)
    if ((s = dalloc(sim, (void**)&p.pv[k], c * 2 * sizeof(float4))) != SC_OK) return s;
  size_t tiles = (c + kBlock - 1) / kBlock;
  if ((s = dalloc(sim, (void**)&p.plans_self, tiles * sizeof(TilePlan))) != SC_OK) return s;
  if ((s = dalloc(sim, (void**)&p.plans_other, tiles * sizeof(TilePlan))) != SC_OK) return s;
  // the four lists' pool counters side by side: one copy samples them all
  if (!sim->pool_ctrs && (s = dalloc(sim, (void**)&sim->pool_ctrs, 4 * kPoolShards * sizeof(uint32_t))) != SC_OK)
    return s;
  for (int q = 0; q < 2; q++) {
    PairList& L = p.lists[q];
    if ((s = dalloc(sim, (void**)&L.e, c * kPairSlots * sizeof(uint2))) != SC_OK) return s;
    if ((s = dalloc(sim, (void**)&L.n, c * sizeof(int))) != SC_OK) return s;
    if ((s = dalloc(sim, (void**)&L.pool_off, c * sizeof(uint32_t))) != SC_OK) return s;
    L.pool_ctr = sim->pool_ctrs + ((metabolic ? 2 : 0) + q) * kPoolShards;
    L.shard_cap = (uint32_t)((c * kPoolFactor + kPoolShards - 1) / kPoolShards);
    if ((s = dalloc(sim, (void**)&L.pool, (size_t)L.shard_cap * kPoolShards * sizeof(uint2))) != SC_OK) return s;
  }
  for (int k = 0; k < 2; k++) {
    if ((s = dalloc(sim, (void**)&p.work[k], tiles * kBlock * sizeof(int))) != SC_OK) return s;
    if ((s = dalloc(sim, (void**)&p.work_cnt[k], tiles * sizeof(int))) != SC_OK) return s;
    if ((s = dalloc(sim, (void**)&p.work_first[k], tiles * sizeof(int))) != SC_OK) return s;
    if ((s = dalloc(sim, (void**)&p.work_bond[k], tiles * sizeof(int))) != SC_OK) return s;
  }
  return SC_OK;
}

template <class T>
static sc_status upload(sc_sim* sim, void* dst, const std::vector<T>& v) {
  if (v.empty()) return SC_OK;
  SC_TRY(mem_copy(sim, dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return SC_OK;
}

// Writes particles into the current buffers of both buffer sets (slots beyond
// n zero), resets cell ranges and sort bookkeeping.
static sc_status upload_pop(sc_sim* sim, Pop& p, const sc_particle* src, size_t n, const int* inst = nullptr) {
  size_t c = (size_t)(p.cap > 0 ? p.cap : 1);
  if (n > (size_t)p.cap) return fail(sim, SC_ERR_BUDGET, "more particles than the array capacity");
  // bonded particles need the bonded relax kernel (the callers drop captured steps)
  if (!p.metabolic)
    for (size_t i = 0; i < n; i++) sim->g.bonds = sim->g.bonds || src[i].n_bonds > 0;
  std::vector<float4> pos(c), vel(c), rot(c), avel(c), pn(c), an(c), dbg(c), m0, m1;
  std::vector<int4> hdr(c), st(c), bonds;
  std::vector<float2> m2;
  std::memset(pos.data(), 0, c * sizeof(float4)); std::memset(vel.data(), 0, c * sizeof(float4));
  std::memset(rot.data(), 0, c * sizeof(float4)); std::memset(avel.data(), 0, c * sizeof(float4));
  std::memset(pn.data(), 0, c * sizeof(float4)); std::memset(an.data(), 0, c * sizeof(float4));
  std::memset(dbg.data(), 0, c * sizeof(float4));
  std::memset(hdr.data(), 0, c * sizeof(int4)); std::memset(st.data(), 0, c * sizeof(int4));
  std::vector<int> ins(c, 0);
  if (inst) std::copy(inst, inst + n, ins.begin());
  if (p.metabolic) {
    m0.assign(c, make_float4(0, 0, 0, 0)); m1.assign(c, make_float4(0, 0, 0, 0)); m2.assign(c, make_float2(0, 0));
  } else {
    bonds.assign(c * SC_MAX_BONDS, make_int4(0, 0, 0, 0));
  }
  for (size_t i = 0; i < n; i++) {
    const sc_particle& s = src[i];
    pos[i] = make_float4(s.pos[0], s.pos[1], s.pos[2], s.radius);
    vel[i] = make_float4(s.vel[0], s.vel[1], s.vel[2], 0.0f);
    rot[i] = make_float4(s.rot[0], s.rot[1], s.rot[2], s.rot[3]);
    avel[i] = make_float4(s.ang_vel[0], s.ang_vel[1], s.ang_vel[2], s.ang_vel[3]);
    pn[i] = make_float4(s.pos_noise[0], s.pos_noise[1], s.pos_noise[2], 0.0f);
    an[i] = make_float4(s.ang_noise[0], s.ang_noise[1], s.ang_noise[2], s.ang_noise[3]);
    dbg[i] = make_float4(s.debug[0], s.debug[1], s.debug[2], s.debug[3]);
    hdr[i] = make_int4(s.id, s.type, s.flags, (int8_t)s.hydrophobic);
    int nb = s.n_bonds;
    if (nb < 0 || nb > SC_MAX_BONDS) return fail(sim, SC_ERR_INVALID, "n_bonds out of range");
    st[i] = make_int4((int8_t)s.state, s.last_state_change_step, nb, s.scheduled_polymerization);
    if (p.metabolic) {
      m0[i] = make_float4(s.metabolites[0], s.metabolites[1], s.metabolites[2], s.metabolites[3]);
      m1[i] = make_float4(s.metabolites[4], s.metabolites[5], s.metabolites[6], s.metabolites[7]);
      m2[i] = make_float2(s.metabolites[8], s.metabolites[9]);
    } else {
      for (int k = 0; k < SC_MAX_BONDS; k++)
        bonds[(size_t)k * c + i] = make_int4(s.bonds[k][0], s.bonds[k][1], s.bonds[k][2], 0);
    }
  }
  for (int f = 0; f < F_COUNT; f++) p.cur[f] = 0;
  for (int k = 0; k < 2; k++) {
    sc_status s;
    if ((s = upload(sim, p.buf[F_POS][k], pos)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_VEL][k], vel)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_ROT][k], rot)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_AVEL][k], avel)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_PNOISE][k], pn)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_ANOISE][k], an)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_DBG][k], dbg)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_HDR][k], hdr)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_ST][k], st)) != SC_OK) return s;
    if ((s = upload(sim, p.buf[F_INST][k], ins)) != SC_OK) return s;
    if (p.metabolic) {
      if ((s = upload(sim, p.buf[F_MET0][k], m0)) != SC_OK) return s;
      if ((s = upload(sim, p.buf[F_MET1][k], m1)) != SC_OK) return s;
      if ((s = upload(sim, p.buf[F_MET2][k], m2)) != SC_OK) return s;
    } else {
      if ((s = upload(sim, p.buf[F_BONDS][k], bonds)) != SC_OK) return s;
    }
  }
  size_t cells = (size_t)sim->g.sentinel;
  SC_TRY(mem_fill(sim, p.cs, 0, (cells + 1) * sizeof(uint32_t)));
  return SC_OK;
}

static sc_status set_counts(sc_sim* sim, const sc_counts& c) {
  Counters h;
  std::memset(&h, 0, sizeof(h));
  h.n_active[0] = c.n_active;
  h.n_active[1] = c.n_active_metabolic;
  h.next_particle_id = c.next_particle_id;
  h.step = c.step;
  SC_TRY(mem_copy(sim, sim->ctr, &h, sizeof(h), hipMemcpyHostToDevice));
  return SC_OK;
}

static sc_status get_counters(sc_sim* sim, Counters& h) {
  SC_TRY(hipStreamSynchronize(sim->stream));
  SC_TRY(mem_copy(sim, &h, sim->ctr, sizeof(h), hipMemcpyDeviceToHost));
  return SC_OK;
}

static void drop_graphs(sc_sim* sim) {
  for (auto& kv : sim->graphs) (void)hipGraphExecDestroy(kv.second.exec);
  sim->graphs.clear();
}

// ---- stage enqueue ------------------------------------------------------------

static void enqueue_stage(sc_sim* sim, int stage, int arg) {
  hipStream_t s = sim->stream;
  Pop& P = sim->pop[0];
  Pop& M = sim->pop[1];
  switch (stage) {
    case SC_STAGE_SORT:
      for (int k = 0; k < 2; k++) {
        launch_keys(sim->pop[k], k, sim->g, sim->ctr, s);
        launch_sort_passes(sim->pop[k], k, sim->g, sim->ctr, s);
        launch_permute(sim->pop[k], k, sim->g, sim->ctr, s);
      }
      launch_tile_plans(P, M, sim->g, sim->ctr, s);
      break;
    case SC_STAGE_MOVE:
      launch_move(P, 0, sim->g, sim->ph, sim->ctr, s);
      launch_move(M, 1, sim->g, sim->ph, sim->ctr, s);
      break;
    case SC_STAGE_COORDINATE:
      launch_coordinate(P, sim->g, sim->ctr, arg, (float)arg / (float)sim->host.P.noise_coordination_steps, s);
      break;
    case SC_STAGE_APPLY_NOISE:
      launch_apply_noise(P, 0, sim->ctr, s);
      launch_apply_noise(M, 1, sim->ctr, s);
      break;
    case SC_STAGE_RELAX:
      if (sim->fusing && arg == 0)
        launch_relax_fused(P, M, sim->g, sim->t, sim->ctr, sim->fuse, sim->host.P.relaxation_steps, s);
      launch_relax(P, sim->g, sim->t, sim->ctr, arg, s, sim->fusing, sim->relax_fork);
      break;
    case SC_STAGE_RELAX_METABOLIC: launch_relax_metabolic(M, P, sim->g, sim->ctr, arg / 2, s, sim->fusing); break;
    case SC_STAGE_RELAX_PARTNERS: launch_relax_partners(P, M, sim->g, sim->ctr, arg == 0, s, sim->fusing); break;
    case SC_STAGE_APPLY_VELOCITIES:
      launch_apply_velocities(P, 0, sim->g, sim->ph, sim->ctr, s);
      launch_apply_velocities(M, 1, sim->g, sim->ph, sim->ctr, s);
      break;
    case SC_STAGE_COMPLEXIFY:
      launch_complexify(P, sim->g, sim->t, sim->ctr, arg, s);
      if (sim->site_rows) launch_site_bonds(P, sim->g, sim->t, sim->ctr, arg, s);
      break;
    case SC_STAGE_TRANSITION: launch_transition(P, sim->g, sim->t, sim->ctr, s); break;
    case SC_STAGE_POLYMERIZE: launch_polymerize(P, sim->g, sim->t, sim->ph, sim->ctr, sim->d_ictr, s); break;
    case SC_STAGE_DIFFUSE: launch_diffuse(M, sim->g, sim->ctr, arg, s); break;
    case SC_STAGE_REMOVE_INTERFERING: launch_remove_interfering(M, P, sim->g, sim->ctr, s); break;
    case SC_STAGE_END_STEP: launch_end_step(sim->ctr, s); break;
    case SC_STAGE_PROPAGATE_STATE: launch_propagate_state(P, sim->g, sim->t, sim->ctr, s); break;
    case SC_STAGE_REACT: launch_react(M, P, sim->g, sim->t, sim->ctr, s); break;
    default: break;
  }
}

// Events for the fork/join points of one concurrent step launched eagerly
// (created outside any capture).
static sc_status ensure_fork_events(sc_sim* sim) {
  const size_t need = 2 * (size_t)sim->host.P.relaxation_steps + 8;
  while (sim->fork_ev.size() < need) {
    hipEvent_t e;
    SC_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    sim->fork_ev.push_back(e);
  }
  return SC_OK;
}

// Fork/join within a step.  Launched eagerly, the side work runs on the
// second stream behind an event and the main stream waits for it through
// another.  Captured, everything is launched on the capturing stream in
// order (the capture then is a chain of nodes, one per launch, and no other
// stream or capture call takes part in it), and the launch numbers of each
// fork, end of side work and join are noted; split_branches then turns the
// noted stretches of the chain into parallel branches.
struct StepFork : Fork {
  struct Branch { uint64_t fork, side_end, join; };
  sc_sim* sim;
  bool capture;
  size_t ev = 0;
  uint64_t base = g_launches;
  Branch cur{0, 0, 0};
  std::vector<Branch> branches;
  StepFork(sc_sim* s, bool c) : sim(s), capture(c) {}
  hipStream_t begin() override {
    if (capture) { cur.fork = g_launches - base; return sim->stream; }
    hipEventRecord(sim->fork_ev[ev], sim->stream);
    hipStreamWaitEvent(sim->side, sim->fork_ev[ev], 0);
    ev++;
    return sim->side;
  }
  void end() override {
    if (capture) cur.side_end = g_launches - base;
  }
  void join() override {
    if (capture) {
      cur.join = g_launches - base;
      branches.push_back(cur);
      return;
    }
    hipEventRecord(sim->fork_ev[ev], sim->side);
    hipStreamWaitEvent(sim->stream, sim->fork_ev[ev], 0);
    ev++;
  }
};

// Rewires a captured chain of `total` nodes: for each branch, the main
// stretch [side_end, join) gets the dependencies of the fork's first node
// instead of the side stretch's last node, and the join's node also waits for
// the side stretch's last node.  The graph is left as captured (still
// correct, only serial) if it is not the expected chain.
static void split_branches(hipGraph_t graph, const std::vector<StepFork::Branch>& branches, uint64_t total) {
  size_t n = 0;
  if (hipGraphGetNodes(graph, nullptr, &n) != hipSuccess || n != total || n == 0) return;
  hipGraphNode_t node;
  size_t k = 1;
  if (hipGraphGetRootNodes(graph, &node, &k) != hipSuccess || k != 1) return;
  std::vector<hipGraphNode_t> chain{node};
  for (;;) {
    size_t d = 0;
    if (hipGraphNodeGetDependentNodes(chain.back(), nullptr, &d) != hipSuccess || d > 1) return;
    if (d == 0) break;
    if (hipGraphNodeGetDependentNodes(chain.back(), &node, &d) != hipSuccess) return;
    chain.push_back(node);
  }
  if (chain.size() != n) return;
  for (const StepFork::Branch& b : branches) {
    if (!(b.fork < b.side_end && b.side_end < b.join && b.join <= n)) continue;
    size_t nd = 0;
    if (hipGraphNodeGetDependencies(chain[b.fork], nullptr, &nd) != hipSuccess) return;
    std::vector<hipGraphNode_t> deps(nd);
    if (nd && hipGraphNodeGetDependencies(chain[b.fork], deps.data(), &nd) != hipSuccess) return;
    hipGraphNode_t side_tail = chain[b.side_end - 1], main_head = chain[b.side_end];
    if (hipGraphRemoveDependencies(graph, &side_tail, &main_head, 1) != hipSuccess) return;
    for (hipGraphNode_t d : deps)
      if (hipGraphAddDependencies(graph, &d, &main_head, 1) != hipSuccess) return;
    if (b.join < n && hipGraphAddDependencies(graph, &side_tail, &chain[b.join], 1) != hipSuccess) return;
  }
}

// main.cu:347-649 with the launches that touch only one population beside
// the others: the metabolic sort, move and noise beside the plain ones and
// the coordination passes; the metabolic velocity update and diffusion beside
// the plain velocity update, complexification, transitions and
// polymerisation; the reactions (which read both populations) after the join.
// Each side run reads and writes only its own population's arrays and
// counters, and is joined before anything that reads them.
static void enqueue_step_concurrent(sc_sim* sim, StepFork& f) {
  const sc_params& P = sim->host.P;
  hipStream_t s = sim->stream;
  Pop& Pp = sim->pop[0];
  Pop& M = sim->pop[1];
  hipStream_t s2 = f.begin();
  launch_keys(M, 1, sim->g, sim->ctr, s2);
  launch_sort_passes(M, 1, sim->g, sim->ctr, s2);
  launch_permute(M, 1, sim->g, sim->ctr, s2);
  f.end();
  launch_keys(Pp, 0, sim->g, sim->ctr, s);
  launch_sort_passes(Pp, 0, sim->g, sim->ctr, s);
  launch_permute(Pp, 0, sim->g, sim->ctr, s);
  f.join();
  launch_tile_plans(Pp, M, sim->g, sim->ctr, s);
  s2 = f.begin();
  launch_move(M, 1, sim->g, sim->ph, sim->ctr, s2);
  launch_apply_noise(M, 1, sim->ctr, s2);
  f.end();
  launch_move(Pp, 0, sim->g, sim->ph, sim->ctr, s);
  for (int j = 0; j < P.noise_coordination_steps; j++)
    launch_coordinate(Pp, sim->g, sim->ctr, j, (float)j / (float)P.noise_coordination_steps, s);
  launch_apply_noise(Pp, 0, sim->ctr, s);
  f.join();
  for (int j = 0; j < P.relaxation_steps; j++) {
    if (sim->fuse_on && j == 0) launch_relax_fused(Pp, M, sim->g, sim->t, sim->ctr, sim->fuse, P.relaxation_steps, s);
    launch_relax(Pp, sim->g, sim->t, sim->ctr, j, s, sim->fuse_on);
    if (j % 2 == 0) {
      launch_relax_metabolic(M, Pp, sim->g, sim->ctr, j / 2, s, sim->fuse_on);
      launch_relax_partners(Pp, M, sim->g, sim->ctr, j == 0, s, sim->fuse_on);
    }
  }
  s2 = f.begin();
  launch_apply_velocities(M, 1, sim->g, sim->ph, sim->ctr, s2);
  for (int j = 0; j < P.metabolite_diffusion_steps; j++) launch_diffuse(M, sim->g, sim->ctr, j, s2);
  f.end();
  launch_apply_velocities(Pp, 0, sim->g, sim->ph, sim->ctr, s);
  launch_complexify(Pp, sim->g, sim->t, sim->ctr, 0, s);
  if (sim->site_rows) launch_site_bonds(Pp, sim->g, sim->t, sim->ctr, 0, s);
  launch_transition(Pp, sim->g, sim->t, sim->ctr, s);
  launch_polymerize(Pp, sim->g, sim->t, sim->ph, sim->ctr, sim->d_ictr, s);
  if (P.propagate_state) launch_propagate_state(Pp, sim->g, sim->t, sim->ctr, s);
  f.join();
  if (sim->t.n_rx > 0) launch_react(M, Pp, sim->g, sim->t, sim->ctr, s);   // reads both populations
  launch_end_step(sim->ctr, s);
}

// main.cu:347-649 as a stage sequence; `hook(stage)` brackets each stage (profiling).
template <class H>
static void enqueue_step(sc_sim* sim, H&& hook) {
  const sc_params& P = sim->host.P;
  auto run = [&](int st, int arg) { hook(st, true); enqueue_stage(sim, st, arg); hook(st, false); };
  struct Fusing {   // whole steps fuse; sc_run_stage's single stages never do
    sc_sim* s;
    Fusing(sc_sim* x) : s(x) { s->fusing = s->fuse_on; }
    ~Fusing() { s->fusing = false; }
  } fusing_scope(sim);
  run(SC_STAGE_SORT, 0);
  run(SC_STAGE_MOVE, 0);
  for (int j = 0; j < P.noise_coordination_steps; j++) run(SC_STAGE_COORDINATE, j);
  run(SC_STAGE_APPLY_NOISE, 0);
  for (int j = 0; j < P.relaxation_steps; j++) {
    run(SC_STAGE_RELAX, j);
    if (j % 2 == 0) { run(SC_STAGE_RELAX_METABOLIC, j); run(SC_STAGE_RELAX_PARTNERS, j); }
  }
  run(SC_STAGE_APPLY_VELOCITIES, 0);
  run(SC_STAGE_COMPLEXIFY, 0);
  run(SC_STAGE_TRANSITION, 0);
  run(SC_STAGE_POLYMERIZE, 0);
  if (P.propagate_state) run(SC_STAGE_PROPAGATE_STATE, 0);
  for (int j = 0; j < P.metabolite_diffusion_steps; j++) run(SC_STAGE_DIFFUSE, j);
  if (sim->t.n_rx > 0) run(SC_STAGE_REACT, 0);
  run(SC_STAGE_END_STEP, 0);
}

// Buffer parities (one bit per field and array) and the grid sizes (launch
// bounds) a captured step depends on.
static uint64_t signature(const sc_sim* sim) {
  uint64_t sig = 0;
  int b = 0;
  for (int k = 0; k < 2; k++)
    for (int f = 0; f < F_COUNT; f++) sig |= (uint64_t)(sim->pop[k].cur[f] & 1) << (b++);
  for (int k = 0; k < 2; k++) {
    int lg = 0;
    while ((1ll << lg) < (long long)sim->pop[k].launch) lg++;
    sig |= (uint64_t)lg << (2 * F_COUNT + 6 * k);
  }
  return sig;
}
static void apply_signature(sc_sim* sim, uint64_t sig) {
  int b = 0;
  for (int k = 0; k < 2; k++)
    for (int f = 0; f < F_COUNT; f++) sim->pop[k].cur[f] = (int)((sig >> (b++)) & 1);
}

// The sequential schedule of a whole step, with the relax passes' two
// kernels (relax_split) beside each other: they touch disjoint particles and
// read only what the pass before them wrote.
static void enqueue_step_forked(sc_sim* sim, StepFork& f) {
  sim->relax_fork = sim->g.relax_split ? &f : nullptr;
  enqueue_step(sim, [](int, bool) {});
  sim->relax_fork = nullptr;
}

static sc_status step_once(sc_sim* sim) {
  if (!sim->use_graph) {
    sc_status st = ensure_fork_events(sim);
    if (st != SC_OK) return st;
    StepFork f(sim, false);
    if (sim->concurrent) enqueue_step_concurrent(sim, f);
    else enqueue_step_forked(sim, f);
    SC_TRY(hipGetLastError());
    return SC_OK;
  }
  uint64_t sig = signature(sim);
  auto it = sim->graphs.find(sig);
  if (it == sim->graphs.end()) {
    hipGraph_t graph = nullptr;
    StepFork f(sim, true);
    SC_TRY(hipStreamBeginCapture(sim->stream, hipStreamCaptureModeThreadLocal));
    if (sim->concurrent) enqueue_step_concurrent(sim, f);
    else enqueue_step_forked(sim, f);
    hipError_t e = hipStreamEndCapture(sim->stream, &graph);
    if (e != hipSuccess) { apply_signature(sim, sig); return fail(sim, SC_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e)); }
    if (!f.branches.empty()) split_branches(graph, f.branches, g_launches - f.base);
    GraphEntry ge;
    e = hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) { apply_signature(sim, sig); return fail(sim, SC_ERR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e)); }
    ge.end_sig = signature(sim);
    apply_signature(sim, sig);
    it = sim->graphs.emplace(sig, ge).first;
  }
  SC_TRY(hipGraphLaunch(it->second.exec, sim->stream));
  apply_signature(sim, it->second.end_sig);
  return SC_OK;
}

// ---- list pool sizing ------------------------------------------------------------
// The list builders count, per pool shard, the entries particles asked for
// past their slots (also those that did not fit: those particles walk the
// stencil in every pass, correct but slow).  After a step the counters are
// copied to pinned memory; before a later step, once that copy has landed, a
// shard that was short grows the pool, and the captured graphs (which hold
// the pool's address) are dropped.
static PairList& list_at(sc_sim* sim, int q) { return sim->pop[q / 2].lists[q % 2]; }

static sc_status pool_sample(sc_sim* sim) {
  if (sim->demand_pending) return SC_OK;
  if (!sim->h_demand) {
    SC_TRY(hipHostMalloc((void**)&sim->h_demand, 4 * kPoolShards * sizeof(uint32_t)));
    SC_TRY(hipEventCreateWithFlags(&sim->demand_ev, hipEventDisableTiming));
  }
  SC_TRY(hipMemcpyAsync(sim->h_demand, sim->pool_ctrs, 4 * kPoolShards * sizeof(uint32_t), hipMemcpyDeviceToHost,
                        sim->stream));   // list q's counters at q * kPoolShards
  SC_TRY(hipEventRecord(sim->demand_ev, sim->stream));
  sim->demand_pending = true;
  return SC_OK;
}

static sc_status pool_check(sc_sim* sim) {
  if (!sim->demand_pending || hipEventQuery(sim->demand_ev) != hipSuccess) return SC_OK;
  sim->demand_pending = false;
  for (int q = 0; q < 4; q++) {
    PairList& L = list_at(sim, q);
    uint32_t most = 0;
    for (int k = 0; k < kPoolShards; k++) most = std::max(most, sim->h_demand[q * kPoolShards + k]);
    // grow before a shard runs full: the demand of a step varies with the
    // builders' per-wave reservations (estimated from the last step's lists)
    if ((uint64_t)most * 5 <= (uint64_t)L.shard_cap * 4) continue;
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    const uint64_t cap = (uint64_t)most + most / 2 + 64;
    if (cap * kPoolShards >= ((uint64_t)1 << 32)) continue;            // offsets are 32-bit: keep walking
    void* grown = nullptr;
    sc_status s = dalloc(sim, &grown, (size_t)cap * kPoolShards * sizeof(uint2));
    if (s != SC_OK) return s;
    auto it = std::find(sim->allocs.begin(), sim->allocs.end(), (void*)L.pool);
    if (it != sim->allocs.end()) sim->allocs.erase(it);
    SC_TRY(hipFree(L.pool));
    L.pool = (uint2*)grown;
    L.shard_cap = (uint32_t)cap;
  }
  return SC_OK;
}

// One step with the pool bookkeeping around it: grow from an earlier
// step's demand if that has landed, sample this step's.
static sc_status step_pooled(sc_sim* sim) {
  sc_status s = pool_check(sim);
  if (s != SC_OK) return s;
  if ((s = step_once(sim)) != SC_OK) return s;
  return pool_sample(sim);
}

// init sequence of main.cu:283-338 (frame persistence is the caller's)
static sc_status run_init(sc_sim* sim) {
  enqueue_stage(sim, SC_STAGE_SORT, 0);
  enqueue_stage(sim, SC_STAGE_REMOVE_INTERFERING, 0);
  enqueue_stage(sim, SC_STAGE_COMPLEXIFY, -1);
  SC_TRY(hipGetLastError());
  SC_TRY(hipStreamSynchronize(sim->stream));
  return SC_OK;
}

template <class T>
static sc_status download(sc_sim* sim, std::vector<T>& v, const void* src, size_t n) {
  v.resize(n);
  if (n) SC_TRY(mem_copy(sim, v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost));
  return SC_OK;
}

extern "C" {

// Instances of a batch share the per-step schedule, the physics constants and
// the tables; they differ in box, seed and populations.
static bool same_shared_constants(const scs::HostSetup& a, const scs::HostSetup& b, std::string& why) {
  const sc_params &P = a.P, &Q = b.P;
  auto same_bytes = [](const auto& x, const auto& y) {
    return x.size() == y.size() && (x.empty() || std::memcmp(x.data(), y.data(), x.size() * sizeof(x[0])) == 0);
  };
  if (P.grid_cell_size != Q.grid_cell_size || P.max_interaction_distance != Q.max_interaction_distance ||
      P.max_diffusion_distance != Q.max_diffusion_distance || P.boundary != Q.boundary) why = "grid constants";
  else if (P.movement_noise_scale != Q.movement_noise_scale || P.rotation_noise_scale != Q.rotation_noise_scale ||
           P.metabolite_movement_noise_scale != Q.metabolite_movement_noise_scale ||
           P.velocity_decay != Q.velocity_decay || P.angular_velocity_decay != Q.angular_velocity_decay ||
           P.noise_model != Q.noise_model || P.se_translation != Q.se_translation ||
           P.se_rotation != Q.se_rotation) why = "noise or decay";
  else if (P.noise_coordination_steps != Q.noise_coordination_steps || P.relaxation_steps != Q.relaxation_steps ||
           P.metabolite_diffusion_steps != Q.metabolite_diffusion_steps) why = "sub-iteration counts";
  else if (!same_bytes(a.flat, b.flat) || !same_bytes(a.partner, b.partner) ||
           !same_bytes(a.partner_index, b.partner_index) || !same_bytes(a.types, b.types)) why = "type or interaction tables";
  else if (!same_bytes(a.reactions, b.reactions)) why = "reactions";
  else if (P.propagate_state != Q.propagate_state) why = "propagateState";
  else if (a.run_init != b.run_init) why = "init";
  else return true;
  return false;
}

static GridP instance_grid(const scs::HostSetup& H) {
  const sc_params& P = H.P;
  GridP g;
  std::memset(&g, 0, sizeof(g));
  g.cell = P.grid_cell_size; g.L = P.sim_size; g.cutoff = P.max_interaction_distance;
  g.diff_cutoff = P.max_diffusion_distance; g.n_cells = H.n_grid_cells; g.bits = H.n_grid_bits;
  g.boundary = P.boundary; g.n_per = H.n_periodic;
  g.seed_lo = P.seed_lo; g.seed_hi = P.seed_hi;
  g.xcd_run = kXcdRunDefault;
  g.coll_gather = kCollGatherDefault;
  g.pass_waves = kPassWavesDefault;
  g.relax_split = 0;
  g.sort_items = kSortItemsDefault;
  // bonds form only through interaction rows (complexify, site contact,
  // polymerisation); otherwise only the particles created with bonds have any
  g.bonds = !H.partner.empty();
  for (const sc_particle& r : H.plain) g.bonds = g.bonds || r.n_bonds > 0;
  return g;
}

// Type and interaction tables of H on the device.
static sc_status upload_tables(sc_sim* sim, const scs::HostSetup& H) {
  sc_status s;
  Tables& t = sim->t;
  void* d;
  if ((s = dalloc(sim, &d, H.flat.size() * sizeof(sc_interaction_info))) != SC_OK) return s;
  if ((s = upload(sim, d, H.flat)) != SC_OK) return s;
  t.flat = (const sc_interaction_info*)d; t.n_flat = (int)H.flat.size();
  if ((s = dalloc(sim, &d, H.partner.size() * sizeof(sc_interaction_info))) != SC_OK) return s;
  if ((s = upload(sim, d, H.partner)) != SC_OK) return s;
  t.partner = (const sc_interaction_info*)d; t.n_partner = (int)H.partner.size();
  if ((s = dalloc(sim, &d, H.partner_index.size() * sizeof(int32_t))) != SC_OK) return s;
  if ((s = upload(sim, d, H.partner_index)) != SC_OK) return s;
  t.pidx = (const int32_t*)d; t.n_pidx = (int)H.partner_index.size();
  if ((s = dalloc(sim, &d, H.types.size() * sizeof(sc_type_info))) != SC_OK) return s;
  if ((s = upload(sim, d, H.types)) != SC_OK) return s;
  t.types = (const sc_type_info*)d; t.n_types = (int)H.types.size();
  if ((s = dalloc(sim, &d, H.reactions.size() * sizeof(sc_reaction))) != SC_OK) return s;
  if ((s = upload(sim, d, H.reactions)) != SC_OK) return s;
  t.rx = (const sc_reaction*)d; t.n_rx = (int)H.reactions.size();
  t.rx_catalysed = 0;
  for (const sc_reaction& r : H.reactions) t.rx_catalysed |= r.catalyst_type >= 0;
  sim->site_rows = false;
  for (const sc_interaction_info& r : H.partner) sim->site_rows = sim->site_rows || r.bond_probability > 0.0f;
  return SC_OK;
}

// Device state of sim->inst (one setup: a single cell; several: a batch).
static sc_status build_sim(sc_sim* sim, int device) {
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(sim, SC_ERR_DEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(sim, SC_ERR_INVALID, "device ordinal out of range");
  sim->device = device;
  if (hipSetDevice(device) != hipSuccess) return fail(sim, SC_ERR_DEVICE, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking) != hipSuccess) return fail(sim, SC_ERR_DEVICE, "stream");
  if (hipStreamCreateWithFlags(&sim->side, hipStreamNonBlocking) != hipSuccess) return fail(sim, SC_ERR_DEVICE, "stream");
  const int ninst = (int)sim->inst.size();
  sim->host = sim->inst[0];
  const scs::HostSetup& H = sim->host;
  const sc_params& P = H.P;
  // key space: a single cell uses the reference's keys; a batch gives each
  // instance the range [base, base + cells) with base aligned to its cell count
  GridP& g = sim->g;
  g = instance_grid(H);
  uint64_t total = 0;
  sim->table.clear();
  for (int i = 0; i < ninst; i++) {
    GridP gi = instance_grid(sim->inst[i]);
    g.bonds = g.bonds || gi.bonds;
    gi.fresh = 1;
    const uint64_t cells = (uint64_t)1 << (3 * gi.bits);
    total = (total + cells - 1) / cells * cells;
    gi.key_base = (uint32_t)total;
    total += cells;
    if (total >= 0xffffffffull) return fail(sim, SC_ERR_BUDGET, "the batch's cells exceed 32-bit keys");
    sim->table.push_back(gi);
  }
  g.sentinel = (uint32_t)total;
  int key_bits = 1;
  while (key_bits < 32 && ((uint64_t)1 << key_bits) <= total) key_bits++;
  g.key_bits = key_bits;
  g.passes = (g.key_bits + kRadixBits - 1) / kRadixBits;
  for (GridP& gi : sim->table) { gi.sentinel = g.sentinel; gi.key_bits = g.key_bits; gi.passes = g.passes; }
  Phys& ph = sim->ph;
  ph.movement_noise[0] = P.movement_noise_scale; ph.movement_noise[1] = P.metabolite_movement_noise_scale;
  ph.rotation_noise = P.rotation_noise_scale; ph.noise_model = P.noise_model;
  ph.se_translation = P.se_translation; ph.se_rotation = P.se_rotation; ph.velocity_decay = P.velocity_decay;
  ph.angular_velocity_decay = P.angular_velocity_decay; ph.seed_lo = P.seed_lo; ph.seed_hi = P.seed_hi;
  ph.coord_steps = P.noise_coordination_steps; ph.capacity_plain = P.num_particles;
  // capacities and initial slots: the instances' arrays one after another
  size_t cap_p = 0, cap_m = 0;
  std::vector<sc_particle> plain, met;
  std::vector<int> inst_p, inst_m;
  std::vector<InstCtr> ictr(ninst);
  for (int i = 0; i < ninst; i++) {
    const scs::HostSetup& Hi = sim->inst[i];
    cap_p += (size_t)Hi.P.num_particles;
    cap_m += (size_t)Hi.P.num_metabolic_particles;
    plain.insert(plain.end(), Hi.plain.begin(), Hi.plain.end());
    met.insert(met.end(), Hi.metabolic.begin(), Hi.metabolic.end());
    inst_p.insert(inst_p.end(), Hi.plain.size(), i);
    inst_m.insert(inst_m.end(), Hi.metabolic.size(), i);
    ictr[i].next_particle_id = Hi.counts.next_particle_id;
    ictr[i].n_plain = (int)Hi.plain.size();
    ictr[i].cap_plain = Hi.P.num_particles;
    ictr[i].n_metabolic = (int)Hi.metabolic.size();
  }
  if (cap_p >= ((size_t)1 << 30) || cap_m >= ((size_t)1 << 30)) return fail(sim, SC_ERR_BUDGET, "capacity above 2^30 slots");
  size_t cells = (size_t)g.sentinel;
  sc_status s;
  if ((s = alloc_pop(sim, sim->pop[0], (int)cap_p, false, cells)) != SC_OK) return s;
  if ((s = alloc_pop(sim, sim->pop[1], (int)cap_m, true, cells)) != SC_OK) return s;
  if ((s = alloc_fuse(sim)) != SC_OK) return s;
  {
    const size_t mixed = std::max(cap_p, (size_t)1) + std::max(cap_m, (size_t)1);
    if ((s = dalloc(sim, (void**)&sim->pop[1].mix_pos, mixed * sizeof(float4))) != SC_OK) return s;
    if ((s = dalloc(sim, (void**)&sim->pop[1].mix_meta, mixed * sizeof(uint32_t))) != SC_OK) return s;
  }
  if ((s = dalloc(sim, (void**)&sim->ctr, sizeof(Counters))) != SC_OK) return s;
  if (sim->batch) {
    if ((s = dalloc(sim, (void**)&sim->d_table, ninst * sizeof(GridP))) != SC_OK) return s;
    if ((s = upload(sim, sim->d_table, sim->table)) != SC_OK) return s;
    if ((s = dalloc(sim, (void**)&sim->d_ictr, ninst * sizeof(InstCtr))) != SC_OK) return s;
    if ((s = upload(sim, sim->d_ictr, ictr)) != SC_OK) return s;
    g.table = sim->d_table;
  }
  if ((s = upload_tables(sim, H)) != SC_OK) return s;
  sim->frame_bytes = std::max(cap_p, (size_t)1) * SC_REDUCED_PARTICLE_BYTES +
                     std::max(cap_m, (size_t)1) * SC_REDUCED_METABOLIC_BYTES;
  if ((s = dalloc(sim, &sim->d_frame, sim->frame_bytes)) != SC_OK) return s;
  if ((s = upload_pop(sim, sim->pop[0], plain.data(), plain.size(), inst_p.data())) != SC_OK) return s;
  if ((s = upload_pop(sim, sim->pop[1], met.data(), met.size(), inst_m.data())) != SC_OK) return s;
  sc_counts c0 = H.counts;
  c0.n_active = (int)plain.size();
  c0.n_active_metabolic = (int)met.size();
  if ((s = set_counts(sim, c0)) != SC_OK) return s;
  if (H.run_init && (s = run_init(sim)) != SC_OK) return s;
  if (sim->batch) {
    for (GridP& gi : sim->table) gi.fresh = 0;
    if ((s = upload(sim, sim->d_table, sim->table)) != SC_OK) return s;
  }
  return SC_OK;
}

sc_status sc_create(const char* config_json, const char* setup_json, int device, sc_sim** out) {
  if (!out) { g_create_error = "out is null"; return SC_ERR_INVALID; }
  *out = nullptr;
  sc_sim* sim = new sc_sim();
  std::string err;
  sim->inst.resize(1);
  if (!scs::build_setup(config_json, setup_json, sim->inst[0], err)) {
    g_create_error = err;
    delete sim;
    return SC_ERR_INVALID;
  }
  sc_status s = build_sim(sim, device);
  if (s != SC_OK) {
    g_create_error = sim->err;
    sc_destroy(sim);
    return s;
  }
  *out = sim;
  return SC_OK;
}

sc_status sc_batch_create(int n, const char* const* config_json, const char* const* setup_json, int device,
                          sc_sim** out) {
  if (!out) { g_create_error = "out is null"; return SC_ERR_INVALID; }
  *out = nullptr;
  if (n < 1 || !config_json || !setup_json) { g_create_error = "need n >= 1 configs and setups"; return SC_ERR_INVALID; }
  sc_sim* sim = new sc_sim();
  sim->batch = true;
  sim->inst.resize(n);
  for (int i = 0; i < n; i++) {
    std::string err;
    if (!config_json[i] || !setup_json[i] || !scs::build_setup(config_json[i], setup_json[i], sim->inst[i], err)) {
      g_create_error = "instance " + std::to_string(i) + ": " + (err.empty() ? "null document" : err);
      delete sim;
      return SC_ERR_INVALID;
    }
    std::string why;
    if (i > 0 && !same_shared_constants(sim->inst[0], sim->inst[i], why)) {
      g_create_error = "instance " + std::to_string(i) + " differs from instance 0 in " + why +
                       " (instances of a batch share them)";
      delete sim;
      return SC_ERR_INVALID;
    }
  }
  sc_status s = build_sim(sim, device);
  if (s != SC_OK) {
    g_create_error = sim->err;
    sc_destroy(sim);
    return s;
  }
  *out = sim;
  return SC_OK;
}

sc_status sc_destroy(sc_sim* sim) {
  if (!sim) return SC_OK;
  (void)hipSetDevice(sim->device);
  if (sim->stream) (void)hipStreamSynchronize(sim->stream);
  drop_graphs(sim);
  if (sim->h_demand) (void)hipHostFree(sim->h_demand);
  if (sim->demand_ev) (void)hipEventDestroy(sim->demand_ev);
  for (void* p : sim->allocs) (void)hipFree(p);
  if (sim->stream) (void)hipStreamDestroy(sim->stream);
  if (sim->side) (void)hipStreamDestroy(sim->side);
  if (sim->frame_ev) { (void)hipEventSynchronize(sim->frame_ev); (void)hipEventDestroy(sim->frame_ev); }
  if (sim->copy) (void)hipStreamDestroy(sim->copy);
  if (sim->retire_ev) (void)hipEventDestroy(sim->retire_ev);
  if (sim->up_host) (void)hipHostFree(sim->up_host);
  for (hipEvent_t e : sim->fork_ev) (void)hipEventDestroy(e);
  delete sim;
  return SC_OK;
}

const char* sc_last_error(const sc_sim* sim) { return sim ? sim->err.c_str() : g_create_error.c_str(); }

sc_status sc_step(sc_sim* sim, int n) {
  if (!sim) return SC_ERR_INVALID;
  if (n < 0) return fail(sim, SC_ERR_INVALID, "n must be >= 0");
  SC_TRY(hipSetDevice(sim->device));
  for (int i = 0; i < n; i++) {
    sc_status s = step_pooled(sim);
    if (s != SC_OK) return s;
  }
  SC_TRY(hipGetLastError());
  return SC_OK;
}

sc_status sc_run_stage(sc_sim* sim, int stage, int arg) {
  if (!sim) return SC_ERR_INVALID;
  if (stage < 0 || stage >= SC_NUM_STAGES) return fail(sim, SC_ERR_INVALID, "unknown stage");
  SC_TRY(hipSetDevice(sim->device));
  enqueue_stage(sim, stage, arg);
  SC_TRY(hipGetLastError());
  SC_TRY(hipStreamSynchronize(sim->stream));
  return SC_OK;
}

sc_status sc_synchronize(sc_sim* sim) {
  if (!sim) return SC_ERR_INVALID;
  SC_TRY(hipSetDevice(sim->device));
  SC_TRY(hipStreamSynchronize(sim->stream));
  if (sim->frame_ev) SC_TRY(hipEventSynchronize(sim->frame_ev));
  return SC_OK;
}

sc_status sc_counts_get(sc_sim* sim, sc_counts* out) {
  if (!sim || !out) return SC_ERR_INVALID;
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  sc_status s = get_counters(sim, h);
  if (s != SC_OK) return s;
  out->n_active = h.n_active[0];
  out->n_active_metabolic = h.n_active[1];
  out->next_particle_id = h.next_particle_id;
  out->step = h.step;
  return SC_OK;
}

sc_status sc_get_params(const sc_sim* sim, sc_params* out) {
  if (!sim || !out) return SC_ERR_INVALID;
  *out = sim->host.P;
  return SC_OK;
}

sc_status sc_get_tables(const sc_sim* sim, sc_tables* out) {
  if (!sim || !out) return SC_ERR_INVALID;
  const scs::HostSetup& H = sim->host;
  out->n_types = (int32_t)H.types.size(); out->types = H.types.data();
  out->n_flat = (int32_t)H.flat.size(); out->flat = H.flat.data();
  out->n_partner_index = (int32_t)H.partner_index.size(); out->partner_index = H.partner_index.data();
  out->n_partner = (int32_t)H.partner.size(); out->partner = H.partner.data();
  out->n_reactions = (int32_t)H.reactions.size(); out->reactions = H.reactions.data();
  return SC_OK;
}

}  // extern "C"

// sc_particle records [0, n) from SoA field arrays in host memory src[f]
// (bonds k of record i at [k * stride + i]).
static void decode_fields(const void* const* src, size_t stride, bool metabolic, size_t n, sc_particle* dst) {
  const size_t c = stride;
  const float4* pos = (const float4*)src[F_POS];
  const float4* vel = (const float4*)src[F_VEL];
  const float4* rot = (const float4*)src[F_ROT];
  const float4* avel = (const float4*)src[F_AVEL];
  const float4* pn = (const float4*)src[F_PNOISE];
  const float4* an = (const float4*)src[F_ANOISE];
  const float4* dbg = (const float4*)src[F_DBG];
  const int4* hdr = (const int4*)src[F_HDR];
  const int4* st = (const int4*)src[F_ST];
  const float4* m0 = metabolic ? (const float4*)src[F_MET0] : nullptr;
  const float4* m1 = metabolic ? (const float4*)src[F_MET1] : nullptr;
  const float2* m2 = metabolic ? (const float2*)src[F_MET2] : nullptr;
  const int4* bonds = metabolic ? nullptr : (const int4*)src[F_BONDS];
  for (size_t i = 0; i < n; i++) {
    sc_particle& r = dst[i];
    std::memset(&r, 0, sizeof(r));
    r.id = hdr[i].x; r.type = hdr[i].y; r.flags = hdr[i].z; r.hydrophobic = hdr[i].w;
    r.state = st[i].x; r.last_state_change_step = st[i].y; r.n_bonds = st[i].z; r.scheduled_polymerization = st[i].w;
    r.radius = pos[i].w;
    r.pos[0] = pos[i].x; r.pos[1] = pos[i].y; r.pos[2] = pos[i].z;
    r.rot[0] = rot[i].x; r.rot[1] = rot[i].y; r.rot[2] = rot[i].z; r.rot[3] = rot[i].w;
    r.vel[0] = vel[i].x; r.vel[1] = vel[i].y; r.vel[2] = vel[i].z;
    r.pos_noise[0] = pn[i].x; r.pos_noise[1] = pn[i].y; r.pos_noise[2] = pn[i].z;
    r.ang_vel[0] = avel[i].x; r.ang_vel[1] = avel[i].y; r.ang_vel[2] = avel[i].z; r.ang_vel[3] = avel[i].w;
    r.ang_noise[0] = an[i].x; r.ang_noise[1] = an[i].y; r.ang_noise[2] = an[i].z; r.ang_noise[3] = an[i].w;
    r.debug[0] = dbg[i].x; r.debug[1] = dbg[i].y; r.debug[2] = dbg[i].z; r.debug[3] = dbg[i].w;
    if (metabolic) {
      float m[10] = {m0[i].x, m0[i].y, m0[i].z, m0[i].w, m1[i].x, m1[i].y, m1[i].z, m1[i].w, m2[i].x, m2[i].y};
      std::memcpy(r.metabolites, m, sizeof(m));
    } else {
      int nb = std::min(std::max(r.n_bonds, 0), SC_MAX_BONDS);
      for (int k2 = 0; k2 < nb; k2++) {
        int4 b = bonds[(size_t)k2 * c + i];
        r.bonds[k2][0] = b.x; r.bonds[k2][1] = b.y; r.bonds[k2][2] = b.z;
      }
    }
  }
}

// Records [0, n) from SoA field arrays src[f] on the device (bonds k of
// record i at [k * stride + i]), sc_particle layout.
static sc_status download_fields(sc_sim* sim, void* const* src, size_t stride, bool metabolic, size_t n,
                                 sc_particle* dst) {
  if (!dst || n == 0) return SC_OK;
  sc_status s;
  std::vector<char> host[F_COUNT];
  const void* hs[F_COUNT] = {};
  for (int f = 0; f < F_COUNT; f++) {
    if (f == F_INST) continue;
    // bonds: all SC_MAX_BONDS planes of `stride` rows; fields the array lacks have no bytes
    const size_t bytes = f == F_BONDS ? field_bytes(F_BONDS, stride, metabolic) : field_bytes((Field)f, n, metabolic);
    if (bytes == 0) continue;
    if ((s = download(sim, host[f], src[f], bytes)) != SC_OK) return s;
    hs[f] = host[f].data();
  }
  decode_fields(hs, stride, metabolic, n, dst);
  return SC_OK;
}

// Records of slots [0, n) of array k (current buffers).
static sc_status download_records(sc_sim* sim, int k, size_t n, sc_particle* dst) {
  Pop& p = sim->pop[k];
  n = std::min(n, (size_t)p.cap);
  void* src[F_COUNT];
  for (int f = 0; f < F_COUNT; f++) src[f] = p.c<void>((Field)f);
  return download_fields(sim, src, (size_t)(p.cap > 0 ? p.cap : 1), p.metabolic, n, dst);
}

// Records of the given slots of array k, gathered on the device first.
static sc_status download_slots(sc_sim* sim, int k, const std::vector<size_t>& slots, sc_particle* dst) {
  const size_t m = slots.size();
  if (!dst || m == 0) return SC_OK;
  Pop& p = sim->pop[k];
  std::vector<int> idx(slots.begin(), slots.end());
  size_t bytes = m * sizeof(int), off[F_COUNT];
  for (int f = 0; f < F_COUNT; f++) {
    off[f] = bytes;
    bytes += (field_bytes((Field)f, m, p.metabolic) + 255) / 256 * 256;
  }
  // staging kept between calls, grown when short
  if (bytes > sim->gather_cap) {
    void* grown = nullptr;
    sc_status s = dalloc(sim, &grown, bytes);
    if (s != SC_OK) return s;
    if (sim->gather_buf) {
      auto it = std::find(sim->allocs.begin(), sim->allocs.end(), sim->gather_buf);
      if (it != sim->allocs.end()) sim->allocs.erase(it);
      SC_TRY(hipFree(sim->gather_buf));
    }
    sim->gather_buf = grown;
    sim->gather_cap = bytes;
  }
  char* buf = (char*)sim->gather_buf;
  void* out[F_COUNT];
  for (int f = 0; f < F_COUNT; f++) out[f] = buf + off[f];
  if (mem_copy(sim, buf, idx.data(), m * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    return fail(sim, SC_ERR_DEVICE, "gather of an instance's slots failed");
  launch_gather(p, (const int*)buf, (int)m, out, sim->stream);
  // every gathered field in one copy
  std::vector<char> host(bytes - off[0]);
  if (mem_copy(sim, host.data(), buf + off[0], host.size(), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(sim, SC_ERR_DEVICE, "gather of an instance's slots failed");
  const void* hs[F_COUNT];
  for (int f = 0; f < F_COUNT; f++) hs[f] = host.data() + (off[f] - off[0]);
  decode_fields(hs, m, p.metabolic, m, dst);
  return SC_OK;
}

static sc_status batch_only(sc_sim* sim, int instance) {
  if (!sim) return SC_ERR_INVALID;
  if (!sim->batch) return fail(sim, SC_ERR_INVALID, "not a batch (use the single-cell call)");
  if (instance < 0 || instance >= (int)sim->inst.size()) return fail(sim, SC_ERR_INVALID, "instance out of range");
  return SC_OK;
}
static sc_status single_only(sc_sim* sim) {
  if (!sim) return SC_ERR_INVALID;
  if (sim->batch) return fail(sim, SC_ERR_INVALID, "a batch: use the sc_batch_* call with an instance");
  return SC_OK;
}

// Slots of array k that belong to instance i, in slot order.
// The slots of `instance` among the first n_active, in slot order: found on
// the device (launch_instance_slots), only the list comes back.
static sc_status islots_reserve(sc_sim* sim, size_t n_active);

static sc_status instance_slots(sc_sim* sim, int k, int instance, size_t n_active, std::vector<size_t>& out) {
  out.clear();
  if (n_active == 0) return SC_OK;
  const size_t tiles = (n_active + kBlock - 1) / kBlock;
  sc_status rs = islots_reserve(sim, n_active);
  if (rs != SC_OK) return rs;
  int* tile_cnt = sim->islots;
  int* slots = tile_cnt + tiles;
  int* total = slots + n_active;
  launch_instance_slots(sim->pop[k], (int)n_active, instance, tile_cnt, slots, total, sim->stream);
  SC_TRY(hipGetLastError());
  int m = 0;
  SC_TRY(mem_copy(sim, &m, total, sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> list;
  sc_status s = download(sim, list, slots, (size_t)m);
  if (s != SC_OK) return s;
  out.assign(list.begin(), list.end());
  return SC_OK;
}

// scratch of the instance slot search over n_active slots: tile counts, slot
// list, total
static sc_status islots_reserve(sc_sim* sim, size_t n_active) {
  const size_t tiles = (n_active + kBlock - 1) / kBlock;
  const size_t need = tiles + n_active + 1;
  if (need > sim->islots_cap) {
    void* grown = nullptr;
    sc_status s = dalloc(sim, &grown, need * sizeof(int));
    if (s != SC_OK) return s;
    if (sim->islots) {
      auto it = std::find(sim->allocs.begin(), sim->allocs.end(), (void*)sim->islots);
      if (it != sim->allocs.end()) sim->allocs.erase(it);
      SC_TRY(hipFree(sim->islots));
    }
    sim->islots = (int*)grown;
    sim->islots_cap = need;
  }
  return SC_OK;
}

extern "C" {

sc_status sc_snapshot(sc_sim* sim, sc_counts* counts, sc_particle* plain, size_t plain_cap,
                      sc_particle* metabolic, size_t metabolic_cap) {
  sc_status s = single_only(sim);
  if (s != SC_OK) return s;
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  s = get_counters(sim, h);
  if (s != SC_OK) return s;
  if (counts) {
    counts->n_active = h.n_active[0]; counts->n_active_metabolic = h.n_active[1];
    counts->next_particle_id = h.next_particle_id; counts->step = h.step;
  }
  if ((s = download_records(sim, 0, plain_cap, plain)) != SC_OK) return s;
  return download_records(sim, 1, metabolic_cap, metabolic);
}

sc_status sc_batch_size(const sc_sim* sim, int* n) {
  if (!sim || !n) return SC_ERR_INVALID;
  *n = sim->batch ? (int)sim->inst.size() : 1;
  return SC_OK;
}

sc_status sc_batch_counts(sc_sim* sim, int instance, sc_counts* out) {
  sc_status s = batch_only(sim, instance);
  if (s != SC_OK) return s;
  if (!out) return fail(sim, SC_ERR_INVALID, "out is null");
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  if ((s = get_counters(sim, h)) != SC_OK) return s;
  InstCtr ic;
  SC_TRY(mem_copy(sim, &ic, sim->d_ictr + instance, sizeof(ic), hipMemcpyDeviceToHost));
  out->n_active = ic.n_plain;
  out->n_active_metabolic = ic.n_metabolic;
  out->next_particle_id = ic.next_particle_id;
  out->step = h.step - sim->table[instance].step0;
  return SC_OK;
}

sc_status sc_batch_snapshot(sc_sim* sim, int instance, sc_counts* counts, sc_particle* plain, size_t plain_cap,
                            sc_particle* metabolic, size_t metabolic_cap) {
  sc_status s = batch_only(sim, instance);
  if (s != SC_OK) return s;
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  if ((s = get_counters(sim, h)) != SC_OK) return s;
  if (counts && (s = sc_batch_counts(sim, instance, counts)) != SC_OK) return s;
  for (int k = 0; k < 2; k++) {
    sc_particle* dst = k == 0 ? plain : metabolic;
    const size_t want = k == 0 ? plain_cap : metabolic_cap;
    if (!dst || want == 0) continue;
    const size_t n = (size_t)h.n_active[k];
    std::vector<size_t> mine;
    if ((s = instance_slots(sim, k, instance, n, mine)) != SC_OK) return s;
    if (mine.size() > want) mine.resize(want);
    if ((s = download_slots(sim, k, mine, dst)) != SC_OK) return s;
    for (size_t r = mine.size(); r < want; r++) std::memset(&dst[r], 0, sizeof(sc_particle));
  }
  return SC_OK;
}

sc_status sc_restore(sc_sim* sim, const sc_counts* counts, const sc_particle* plain, size_t n_plain,
                     const sc_particle* metabolic, size_t n_metabolic) {
  if (!sim || !counts) return SC_ERR_INVALID;
  if (single_only(sim) != SC_OK) return SC_ERR_INVALID;
  SC_TRY(hipSetDevice(sim->device));
  SC_TRY(hipStreamSynchronize(sim->stream));
  if (counts->n_active < 0 || (size_t)counts->n_active > n_plain ||
      counts->n_active_metabolic < 0 || (size_t)counts->n_active_metabolic > n_metabolic)
    return fail(sim, SC_ERR_INVALID, "counts exceed the records passed");
  sc_status s;
  if ((s = upload_pop(sim, sim->pop[0], plain, n_plain)) != SC_OK) return s;
  if ((s = upload_pop(sim, sim->pop[1], metabolic, n_metabolic)) != SC_OK) return s;
  drop_graphs(sim);
  return set_counts(sim, *counts);
}

static void frame_header(const sc_params& P, void* dst12) {
  unsigned char* o = (unsigned char*)dst12;
  uint32_t a = (uint32_t)P.num_particles, b = (uint32_t)P.num_metabolic_particles;
  std::memcpy(o, &P.sim_size, 4);
  std::memcpy(o + 4, &a, 4);
  std::memcpy(o + 8, &b, 4);
}
static size_t frame_size(const sc_params& P) {
  return 4 + (size_t)P.num_particles * SC_REDUCED_PARTICLE_BYTES + 4 +
         (size_t)P.num_metabolic_particles * SC_REDUCED_METABOLIC_BYTES;
}

sc_status sc_frame_header(const sc_sim* sim, void* dst12) {
  if (!sim || !dst12 || sim->batch) return SC_ERR_INVALID;
  frame_header(sim->host.P, dst12);
  return SC_OK;
}

sc_status sc_frame_size(const sc_sim* sim, size_t* bytes) {
  if (!sim || !bytes || sim->batch) return SC_ERR_INVALID;
  *bytes = frame_size(sim->host.P);
  return SC_OK;
}

sc_status sc_batch_frame_header(const sc_sim* sim, int instance, void* dst12) {
  if (!sim || !dst12 || !sim->batch || instance < 0 || instance >= (int)sim->inst.size()) return SC_ERR_INVALID;
  frame_header(sim->inst[instance].P, dst12);
  return SC_OK;
}

sc_status sc_batch_frame_size(const sc_sim* sim, int instance, size_t* bytes) {
  if (!sim || !bytes || !sim->batch || instance < 0 || instance >= (int)sim->inst.size()) return SC_ERR_INVALID;
  *bytes = frame_size(sim->inst[instance].P);
  return SC_OK;
}

sc_status sc_batch_emit_frame(sc_sim* sim, int instance, void* dst, size_t cap, size_t* written) {
  sc_status s = batch_only(sim, instance);
  if (s != SC_OK) return s;
  const sc_params& Pi = sim->inst[instance].P;
  const size_t need = frame_size(Pi);
  if (written) *written = need;
  if (!dst) return SC_OK;
  if (cap < need) return fail(sim, SC_ERR_INVALID, "frame buffer too small");
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  if ((s = get_counters(sim, h)) != SC_OK) return s;
  // reduce every slot of the batch, keep this instance's records in slot
  // order, zero records up to the instance's capacity (storage.cuh:29-39)
  const size_t cap_p = (size_t)std::max(sim->pop[0].cap, 1);
  unsigned char* dp = (unsigned char*)sim->d_frame;
  unsigned char* dm = dp + cap_p * SC_REDUCED_PARTICLE_BYTES;
  launch_reduce(sim->pop[0], 0, sim->ctr, dp, sim->stream);
  launch_reduce(sim->pop[1], 1, sim->ctr, dm, sim->stream);
  SC_TRY(hipGetLastError());
  SC_TRY(hipStreamSynchronize(sim->stream));
  unsigned char* o = (unsigned char*)dst;
  const int caps[2] = {Pi.num_particles, Pi.num_metabolic_particles};
  const size_t rec[2] = {SC_REDUCED_PARTICLE_BYTES, SC_REDUCED_METABOLIC_BYTES};
  for (int k = 0; k < 2; k++) {
    const size_t n = (size_t)h.n_active[k];
    std::vector<size_t> mine;
    if ((s = instance_slots(sim, k, instance, n, mine)) != SC_OK) return s;
    std::vector<unsigned char> all(n * rec[k]);
    if (n) SC_TRY(mem_copy(sim, all.data(), k == 0 ? dp : dm, n * rec[k], hipMemcpyDeviceToHost));
    const uint32_t c = (uint32_t)caps[k];
    std::memcpy(o, &c, 4);
    o += 4;
    std::memset(o, 0, (size_t)caps[k] * rec[k]);
    for (size_t r = 0; r < mine.size() && r < (size_t)caps[k]; r++) std::memcpy(o + r * rec[k], &all[mine[r] * rec[k]], rec[k]);
    o += (size_t)caps[k] * rec[k];
  }
  return SC_OK;
}

sc_status sc_emit_frame_async(sc_sim* sim, void* dst, size_t cap, size_t* written) {
  if (single_only(sim) != SC_OK) return SC_ERR_INVALID;
  size_t need;
  sc_frame_size(sim, &need);
  if (written) *written = need;
  if (!dst) return SC_OK;
  if (cap < need) return fail(sim, SC_ERR_INVALID, "frame buffer too small");
  SC_TRY(hipSetDevice(sim->device));
  if (!sim->copy) {
    SC_TRY(hipStreamCreateWithFlags(&sim->copy, hipStreamNonBlocking));
    SC_TRY(hipEventCreateWithFlags(&sim->frame_ev, hipEventDisableTiming));
  }
  const sc_params& P = sim->host.P;
  unsigned char* o = (unsigned char*)dst;
  size_t plain_bytes = (size_t)P.num_particles * SC_REDUCED_PARTICLE_BYTES;
  size_t met_bytes = (size_t)P.num_metabolic_particles * SC_REDUCED_METABOLIC_BYTES;
  unsigned char* dp = (unsigned char*)sim->d_frame;
  unsigned char* dm = dp + (size_t)(P.num_particles > 0 ? P.num_particles : 1) * SC_REDUCED_PARTICLE_BYTES;
  uint32_t a = (uint32_t)P.num_particles, b = (uint32_t)P.num_metabolic_particles;
  std::memcpy(o, &a, 4);
  std::memcpy(o + 4 + plain_bytes, &b, 4);
  // the reduce overwrites the device frame: after the previous frame's copy
  SC_TRY(hipStreamWaitEvent(sim->stream, sim->frame_ev, 0));
  if (P.num_particles > 0) launch_reduce(sim->pop[0], 0, sim->ctr, dp, sim->stream);
  if (P.num_metabolic_particles > 0) launch_reduce(sim->pop[1], 1, sim->ctr, dm, sim->stream);
  SC_TRY(hipGetLastError());
  SC_TRY(hipEventRecord(sim->frame_ev, sim->stream));
  SC_TRY(hipStreamWaitEvent(sim->copy, sim->frame_ev, 0));
  if (plain_bytes) SC_TRY(hipMemcpyAsync(o + 4, dp, plain_bytes, hipMemcpyDeviceToHost, sim->copy));
  if (met_bytes) SC_TRY(hipMemcpyAsync(o + 8 + plain_bytes, dm, met_bytes, hipMemcpyDeviceToHost, sim->copy));
  SC_TRY(hipEventRecord(sim->frame_ev, sim->copy));
  return SC_OK;
}

sc_status sc_frame_wait(sc_sim* sim) {
  if (!sim) return SC_ERR_INVALID;
  if (!sim->frame_ev) return SC_OK;
  SC_TRY(hipSetDevice(sim->device));
  SC_TRY(hipEventSynchronize(sim->frame_ev));
  return SC_OK;
}

sc_status sc_host_alloc(size_t bytes, void** out) {
  if (!out) return SC_ERR_INVALID;
  *out = nullptr;
  if (bytes == 0) return SC_OK;
  hipError_t e = hipHostMalloc(out, bytes, hipHostMallocDefault);
  if (e != hipSuccess) { *out = nullptr; return e == hipErrorOutOfMemory ? SC_ERR_OOM : SC_ERR_DEVICE; }
  return SC_OK;
}

sc_status sc_host_free(void* p) {
  if (p && hipHostFree(p) != hipSuccess) return SC_ERR_DEVICE;
  return SC_OK;
}

sc_status sc_emit_frame(sc_sim* sim, void* dst, size_t cap, size_t* written) {
  if (single_only(sim) != SC_OK) return SC_ERR_INVALID;
  size_t need;
  sc_frame_size(sim, &need);
  if (written) *written = need;
  if (!dst) return SC_OK;
  if (cap < need) return fail(sim, SC_ERR_INVALID, "frame buffer too small");
  SC_TRY(hipSetDevice(sim->device));
  const sc_params& P = sim->host.P;
  unsigned char* o = (unsigned char*)dst;
  size_t plain_bytes = (size_t)P.num_particles * SC_REDUCED_PARTICLE_BYTES;
  size_t met_bytes = (size_t)P.num_metabolic_particles * SC_REDUCED_METABOLIC_BYTES;
  unsigned char* dp = (unsigned char*)sim->d_frame;
  unsigned char* dm = dp + (size_t)(P.num_particles > 0 ? P.num_particles : 1) * SC_REDUCED_PARTICLE_BYTES;
  if (sim->frame_ev) SC_TRY(hipEventSynchronize(sim->frame_ev));     // an asynchronous frame's copy
  if (P.num_particles > 0) launch_reduce(sim->pop[0], 0, sim->ctr, dp, sim->stream);
  if (P.num_metabolic_particles > 0) launch_reduce(sim->pop[1], 1, sim->ctr, dm, sim->stream);
  SC_TRY(hipGetLastError());
  SC_TRY(hipStreamSynchronize(sim->stream));
  uint32_t a = (uint32_t)P.num_particles, b = (uint32_t)P.num_metabolic_particles;
  std::memcpy(o, &a, 4);
  if (plain_bytes) SC_TRY(mem_copy(sim, o + 4, dp, plain_bytes, hipMemcpyDeviceToHost));
  std::memcpy(o + 4 + plain_bytes, &b, 4);
  if (met_bytes) SC_TRY(mem_copy(sim, o + 8 + plain_bytes, dm, met_bytes, hipMemcpyDeviceToHost));
  return SC_OK;
}

sc_status sc_debug_export(sc_sim* sim, int what, int population, void* dst, size_t cap, size_t* written) {
  if (single_only(sim) != SC_OK) return SC_ERR_INVALID;
  if (!sim || population < 0 || population > 1) return SC_ERR_INVALID;
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  sc_status s = get_counters(sim, h);
  if (s != SC_OK) return s;
  Pop& p = sim->pop[population];
  const GridP& g = sim->g;
  std::vector<uint32_t> out;
  if (what == SC_EXPORT_KEYS || what == SC_EXPORT_PERMUTATION) {
    int n = h.n_sorted[population];
    if ((s = download(sim, out, what == SC_EXPORT_KEYS ? p.keys_sorted : p.vals_sorted, (size_t)n)) != SC_OK) return s;
    if (what == SC_EXPORT_KEYS)
      for (auto& k : out) if (k == g.sentinel) k = 0xFFFFFFFFu;
  } else if (what == SC_EXPORT_RANGES) {
    size_t cells = (size_t)g.sentinel;
    std::vector<uint32_t> cs;
    if ((s = download(sim, cs, p.cs, cells + 1)) != SC_OK) return s;
    for (size_t c = 0; c < cells; c++)
      if (cs[c + 1] > cs[c]) { out.push_back((uint32_t)c); out.push_back(cs[c]); out.push_back(cs[c + 1]); }
  } else if (what == SC_EXPORT_NEIGHBORS) {
    int n = h.n_active[population];
    void* dcounts = nullptr;
    void* dpairs = nullptr;
    size_t c = (size_t)(p.cap > 0 ? p.cap : 1);
    if (hipMalloc(&dcounts, c * 4) != hipSuccess) return fail(sim, SC_ERR_OOM, "neighbour export");
    launch_neighbor_count(p, population, g, sim->ctr, (uint32_t*)dcounts, sim->stream);
    if (hipStreamSynchronize(sim->stream) != hipSuccess) { (void)hipFree(dcounts); return fail(sim, SC_ERR_DEVICE, "neighbour export"); }
    std::vector<uint32_t> cnt;
    s = download(sim, cnt, dcounts, (size_t)n);
    if (s != SC_OK) { hipFree(dcounts); return s; }
    std::vector<uint32_t> off(n);
    uint64_t total = 0;
    for (int i = 0; i < n; i++) { off[i] = (uint32_t)total; total += cnt[i]; }
    if (n > 0) mem_copy(sim, dcounts, off.data(), (size_t)n * 4, hipMemcpyHostToDevice);
    if (hipMalloc(&dpairs, (size_t)(total > 0 ? total : 1) * 8) != hipSuccess) { hipFree(dcounts); return fail(sim, SC_ERR_OOM, "neighbour export"); }
    launch_neighbor_fill(p, population, g, sim->ctr, (const uint32_t*)dcounts, (int32_t*)dpairs, sim->stream);
    if (hipStreamSynchronize(sim->stream) != hipSuccess) s = fail(sim, SC_ERR_DEVICE, "neighbour export");
    else s = download(sim, out, dpairs, (size_t)total * 2);
    hipFree(dcounts);
    hipFree(dpairs);
    if (s != SC_OK) return s;
  } else if (what == SC_EXPORT_FUSED) {
    int n = h.n_active[population];
    std::vector<uint32_t> n0;
    if ((s = download(sim, n0, p.lists[0].n, (size_t)n)) != SC_OK) return s;
    out.resize((size_t)n);
    for (int i = 0; i < n; i++) out[i] = (n0[i] & kMetaFusedBit) ? 1u : 0u;
  } else if (what == SC_EXPORT_LIST_COUNTS) {
    int n = h.n_active[population];
    std::vector<uint32_t> n0, n1;
    if ((s = download(sim, n0, p.lists[0].n, (size_t)n)) != SC_OK) return s;
    if ((s = download(sim, n1, p.lists[1].n, (size_t)n)) != SC_OK) return s;
    out.resize((size_t)n * 2);
    for (int i = 0; i < n; i++) { out[2 * i] = n0[i] & 0xffffu; out[2 * i + 1] = n1[i] & 0xffffu; }
  } else {
    return fail(sim, SC_ERR_INVALID, "unknown export");
  }
  size_t need = out.size() * 4;
  if (written) *written = need;
  if (dst) {
    if (cap < need) return fail(sim, SC_ERR_INVALID, "export buffer too small");
    std::memcpy(dst, out.data(), need);
  }
  return SC_OK;
}

sc_status sc_profile_steps(sc_sim* sim, int n, double* stage_ms, int n_stages) {
  if (!sim || n < 0 || !stage_ms || n_stages < 0 || n_stages > SC_NUM_STAGES) return SC_ERR_INVALID;
  SC_TRY(hipSetDevice(sim->device));
  std::vector<double> acc(SC_NUM_STAGES, 0.0);
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_stage;
  for (int it = 0; it < n; it++) {
    size_t used = 0;
    auto hook = [&](int st, bool begin) {
      if (used >= ev.size()) { hipEvent_t e; hipEventCreate(&e); ev.push_back(e); }
      hipEventRecord(ev[used], sim->stream);
      if (begin) ev_stage.push_back(st);
      used++;
    };
    ev_stage.clear();
    sc_status s = pool_check(sim);
    if (s != SC_OK) return s;
    enqueue_step(sim, hook);
    SC_TRY(hipGetLastError());
    if ((s = pool_sample(sim)) != SC_OK) return s;
    SC_TRY(hipStreamSynchronize(sim->stream));
    for (size_t k = 0; k + 1 < used; k += 2) {
      float ms = 0.0f;
      hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
      acc[ev_stage[k / 2]] += ms;
    }
  }
  for (hipEvent_t e : ev) hipEventDestroy(e);
  for (int k = 0; k < n_stages; k++) stage_ms[k] = n > 0 ? acc[k] / n : 0.0;
  return SC_OK;
}

sc_status sc_time_steps(sc_sim* sim, int n, double* ms) {
  if (!sim || n < 0 || !ms) return SC_ERR_INVALID;
  SC_TRY(hipSetDevice(sim->device));
  hipEvent_t a, b;
  SC_TRY(hipEventCreate(&a));
  SC_TRY(hipEventCreate(&b));
  SC_TRY(hipEventRecord(a, sim->stream));
  for (int i = 0; i < n; i++) {
    sc_status s = step_pooled(sim);
    if (s != SC_OK) { hipEventDestroy(a); hipEventDestroy(b); return s; }
  }
  SC_TRY(hipEventRecord(b, sim->stream));
  SC_TRY(hipEventSynchronize(b));
  float t = 0.0f;
  hipEventElapsedTime(&t, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  *ms = n > 0 ? (double)t / n : 0.0;
  return SC_OK;
}

sc_status sc_time_kernel(sc_sim* sim, int kernel, int reps, double* ms) {
  if (!sim || !ms || reps < 1 || kernel < SC_KERNEL_RELAX_PASS || kernel > SC_KERNEL_RELAX_FUSED)
    return SC_ERR_INVALID;
  if (sim->host.P.relaxation_steps < 4 || sim->host.P.metabolite_diffusion_steps < 3)
    return fail(sim, SC_ERR_INVALID, "the timed passes need relaxationSteps >= 4 and metaboliteDiffusionSteps >= 3");
  Counters h;
  sc_status st = get_counters(sim, h);
  if (st != SC_OK) return st;
  if (h.step < 1) return fail(sim, SC_ERR_INVALID, "run a step first: the passes use that step's neighbour lists");
  SC_TRY(hipSetDevice(sim->device));
  // the launches queued back to back (the host queues faster than they
  // run), each recording the start and end of its own dispatch in a pair of
  // events: the mean of their elapsed times is the kernel's time on the device
  std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
  sc_status out = SC_OK;
  for (hipEvent_t& e : ev)
    if (hipEventCreate(&e) != hipSuccess) { out = fail(sim, SC_ERR_DEVICE, "event create"); break; }
  if (out == SC_OK) {
    // one untimed launch first (code and data warm)
    launch_pass_only(kernel, sim->pop[0], sim->pop[1], sim->g, sim->t, sim->ctr, sim->fuse,
                     sim->host.P.relaxation_steps, sim->stream);
    for (int r = 0; r < reps; r++)
      launch_pass_only(kernel, sim->pop[0], sim->pop[1], sim->g, sim->t, sim->ctr, sim->fuse,
                       sim->host.P.relaxation_steps, sim->stream, ev[2 * r], ev[2 * r + 1]);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(sim->stream);
    double sum = 0.0;
    for (int r = 0; r < reps && e == hipSuccess; r++) {
      float t = 0.0f;
      e = hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]);
      sum += t;
    }
    if (e != hipSuccess) out = fail(sim, SC_ERR_DEVICE, std::string("timed launches: ") + hipGetErrorString(e));
    else *ms = sum / reps;
  }
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  return out;
}

sc_status sc_time_step_kernel(sc_sim* sim, int kernel, int steps, double* ms) {
  if (!sim || !ms || steps < 1 || kernel != SC_KERNEL_RELAX_PASS) return SC_ERR_INVALID;
  if (sim->concurrent) return fail(sim, SC_ERR_INVALID, "timing in steps needs the sequential schedule (concurrent 0)");
  SC_TRY(hipSetDevice(sim->device));
  SC_TRY(hipStreamSynchronize(sim->stream));
  const int per_step = sim->host.P.relaxation_steps;
  std::vector<hipEvent_t> ev(2 * (size_t)steps * per_step, nullptr);
  sc_status out = SC_OK;
  for (hipEvent_t& e : ev)
    if (hipEventCreate(&e) != hipSuccess) { out = fail(sim, SC_ERR_DEVICE, "event create"); break; }
  PassTimer pt{ev.data(), steps * per_step, 0};
  if (out == SC_OK) {
    // the steps launched one kernel at a time (a captured graph would not
    // carry the events), then the graphs again
    const bool graph = sim->use_graph;
    sim->use_graph = false;
    g_pass_timer = &pt;
    for (int k = 0; k < steps && out == SC_OK; k++) out = step_pooled(sim);
    g_pass_timer = nullptr;
    sim->use_graph = graph;
  }
  if (out == SC_OK) {
    hipError_t e = hipStreamSynchronize(sim->stream);
    double sum = 0.0;
    for (int r = 0; r < pt.n && e == hipSuccess; r++) {
      float t = 0.0f;
      e = hipEventElapsedTime(&t, ev[2 * r], ev[2 * r + 1]);
      sum += t;
    }
    if (e != hipSuccess) out = fail(sim, SC_ERR_DEVICE, std::string("timed steps: ") + hipGetErrorString(e));
    else if (pt.n == 0) out = fail(sim, SC_ERR_INVALID, "no relax pass was launched");
    else *ms = sum / pt.n;
  }
  (void)hipStreamSynchronize(sim->stream);
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  return out;
}

sc_status sc_set_option(sc_sim* sim, const char* name, int value) {
  if (!sim || !name) return SC_ERR_INVALID;
  if (std::strcmp(name, "graph") == 0) { sim->use_graph = value != 0; return SC_OK; }
  if (std::strcmp(name, "concurrent") == 0) {
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    if (value < 0 || value > 2) return fail(sim, SC_ERR_INVALID, "concurrent must be 0, 1 or 2");
    sim->concurrent = value;
    return SC_OK;
  }
  if (std::strcmp(name, "fuse") == 0) {
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    sim->fuse_on = value != 0;
    return SC_OK;
  }
  if (std::strcmp(name, "xcd_run") == 0) {
    if (value < 0) return fail(sim, SC_ERR_INVALID, "xcd_run must be >= 0");
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);                        // the captured launches hold the old value
    sim->g.xcd_run = value;
    return SC_OK;
  }
  if (std::strcmp(name, "relax_split") == 0) {
    if (value != 0 && value != 1) return fail(sim, SC_ERR_INVALID, "relax_split must be 0 or 1");
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    sim->g.relax_split = value;
    return SC_OK;
  }
  if (std::strcmp(name, "fuse_bonds") == 0) {
    if (value != 0 && value != 1) return fail(sim, SC_ERR_INVALID, "fuse_bonds must be 0 or 1");
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    sim->g.fuse_bonds = value;
    return SC_OK;
  }
  if (std::strcmp(name, "sort_items") == 0) {
    if (value != 4 && value != 8 && value != 16) return fail(sim, SC_ERR_INVALID, "sort_items must be 4, 8 or 16");
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    sim->g.sort_items = value;
    return SC_OK;
  }
  if (std::strcmp(name, "pass_waves") == 0) {
    if (value != 3 && value != 4) return fail(sim, SC_ERR_INVALID, "pass_waves must be 3 or 4");
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    sim->g.pass_waves = value;
    return SC_OK;
  }
  if (std::strcmp(name, "coll_gather") == 0) {
    if (value != 4 && value != 8) return fail(sim, SC_ERR_INVALID, "coll_gather must be 4 or 8");
    SC_TRY(hipSetDevice(sim->device));
    SC_TRY(hipStreamSynchronize(sim->stream));
    drop_graphs(sim);
    sim->g.coll_gather = value;
    return SC_OK;
  }
  return fail(sim, SC_ERR_INVALID, std::string("unknown option ") + name);
}

// ---- autotuner ------------------------------------------------------------------
// Picks the fastest setting of the result-invariant scheduling knobs for
// this handle's workload by timing steps.  Every candidate starts from the
// same device-side checkpoint of the whole state (both buffers of every
// field, the counters, a batch's per-instance counters), and the checkpoint
// is restored at the end, so tuning does not advance the simulation and the
// results of later steps are the same whatever is chosen.
struct Checkpoint {
  std::vector<void*> copies;   // [pop][field][buffer], null where a field has no buffer
  int cur[2][F_COUNT];
  Counters ctr;
  std::vector<InstCtr> ictr;
};

static void checkpoint_free(Checkpoint& c) {
  for (void* p : c.copies)
    if (p) (void)hipFree(p);
  c.copies.clear();
}

static sc_status checkpoint_save(sc_sim* sim, Checkpoint& c) {
  c.copies.assign(2 * F_COUNT * 2, nullptr);
  for (int k = 0; k < 2; k++) {
    Pop& p = sim->pop[k];
    for (int f = 0; f < F_COUNT; f++) {
      c.cur[k][f] = p.cur[f];
      const size_t b = field_bytes((Field)f, (size_t)p.cap, p.metabolic);
      for (int q = 0; q < 2; q++) {
        if (!b || !p.buf[f][q]) continue;
        void*& d = c.copies[(k * F_COUNT + f) * 2 + q];
        if (hipMalloc(&d, b) != hipSuccess) { d = nullptr; checkpoint_free(c); return fail(sim, SC_ERR_OOM, "autotune checkpoint"); }
        SC_TRY(hipMemcpyAsync(d, p.buf[f][q], b, hipMemcpyDeviceToDevice, sim->stream));
      }
    }
  }
  sc_status s = get_counters(sim, c.ctr);
  if (s != SC_OK) return s;
  if (sim->batch && sim->d_ictr) {
    c.ictr.resize(sim->table.size());
    SC_TRY(mem_copy(sim, c.ictr.data(), sim->d_ictr, c.ictr.size() * sizeof(InstCtr), hipMemcpyDeviceToHost));
  }
  return SC_OK;
}

static sc_status checkpoint_restore(sc_sim* sim, const Checkpoint& c) {
  SC_TRY(hipStreamSynchronize(sim->stream));
  for (int k = 0; k < 2; k++) {
    Pop& p = sim->pop[k];
    for (int f = 0; f < F_COUNT; f++) {
      p.cur[f] = c.cur[k][f];
      const size_t b = field_bytes((Field)f, (size_t)p.cap, p.metabolic);
      for (int q = 0; q < 2; q++) {
        void* d = c.copies[(k * F_COUNT + f) * 2 + q];
        if (d) SC_TRY(hipMemcpyAsync(p.buf[f][q], d, b, hipMemcpyDeviceToDevice, sim->stream));
      }
    }
  }
  SC_TRY(mem_copy(sim, sim->ctr, &c.ctr, sizeof(Counters), hipMemcpyHostToDevice));
  if (!c.ictr.empty())
    SC_TRY(mem_copy(sim, sim->d_ictr, c.ictr.data(), c.ictr.size() * sizeof(InstCtr), hipMemcpyHostToDevice));
  return SC_OK;
}

sc_status sc_autotune(sc_sim* sim, int steps, char* report, size_t report_cap) {
  if (!sim || steps < 1) return SC_ERR_INVALID;
  if (sim->server) return fail(sim, SC_ERR_STATE, "a serving batch is tuned through its template, not directly");
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  sc_status st = get_counters(sim, h);
  if (st != SC_OK) return st;
  Checkpoint ck;
  if ((st = checkpoint_save(sim, ck)) != SC_OK) return st;
  struct Cand { int split, gather, xcd, items, fuse, waves; double ms; };
  const Cand start{sim->g.relax_split, sim->g.coll_gather, sim->g.xcd_run, sim->g.sort_items, sim->g.fuse_bonds,
                   sim->g.pass_waves, 0.0};
  auto apply = [&](const Cand& c) {
    drop_graphs(sim);
    sim->g.relax_split = c.split;
    sim->g.coll_gather = c.gather;
    sim->g.xcd_run = c.xcd;
    sim->g.sort_items = c.items;
    sim->g.fuse_bonds = c.fuse;
    sim->g.pass_waves = c.waves;
  };
  // one untimed step captures the candidate's graph; then `steps` timed steps
  auto measure = [&](Cand& c) -> sc_status {
    sc_status r = checkpoint_restore(sim, ck);
    if (r != SC_OK) return r;
    apply(c);
    if ((r = step_once(sim)) != SC_OK) return r;
    SC_TRY(hipStreamSynchronize(sim->stream));
    if ((r = checkpoint_restore(sim, ck)) != SC_OK) return r;
    double ms = 0.0;
    if ((r = sc_time_steps(sim, steps, &ms)) != SC_OK) return r;
    c.ms = ms;
    return SC_OK;
  };
  std::vector<Cand> tried;
  Cand best{};
  bool have = false;
  auto consider = [&](Cand c) -> sc_status {
    for (const Cand& t : tried)
      if (t.split == c.split && t.gather == c.gather && t.xcd == c.xcd && t.items == c.items && t.fuse == c.fuse &&
          t.waves == c.waves)
        return SC_OK;
    sc_status r = measure(c);
    if (r != SC_OK) return r;
    tried.push_back(c);
    // a later candidate must beat the best so far by 1 %: candidates that tie
    // within the timing noise keep the earlier (simpler) setting
    if (!have || c.ms < best.ms * 0.99) { best = c; have = true; }
    return SC_OK;
  };
  // the relax kernels first (one kernel at three or four waves per SIMD;
  // two kernels with either gather width), then the tile order of the
  // neighbour kernels, then the sort's tile size, each with the best of the
  // earlier stages
  const Cand stage1[] = {{0, kCollGatherDefault, kXcdRunDefault, kSortItemsDefault, 0, kPassWavesDefault, 0.0},
                         {0, kCollGatherDefault, kXcdRunDefault, kSortItemsDefault, 0, 4, 0.0},
                         {1, 8, kXcdRunDefault, kSortItemsDefault, 0, kPassWavesDefault, 0.0},
                         {1, 4, kXcdRunDefault, kSortItemsDefault, 0, kPassWavesDefault, 0.0}};
  for (const Cand& c : stage1)
    if ((st = consider(c)) != SC_OK) break;
  if (st == SC_OK)
    for (int xcd : {16, 64}) {
      Cand c = best;
      c.xcd = xcd;
      if ((st = consider(c)) != SC_OK) break;
    }
  if (st == SC_OK)
    for (int items : {8, 4}) {
      Cand c = best;
      c.items = items;
      if ((st = consider(c)) != SC_OK) break;
    }
  if (st == SC_OK && sim->g.bonds) {   // bond terms in the pass kernel: only where bonds can exist
    Cand c = best;
    c.fuse = 1;
    st = consider(c);
  }
  sc_status st2 = checkpoint_restore(sim, ck);
  checkpoint_free(ck);
  apply(st == SC_OK && have ? best : start);
  if (st != SC_OK) return st;
  if (st2 != SC_OK) return st2;
  if (report && report_cap) {
    std::string r = "{\"chosen\": {\"relax_split\": " + std::to_string(best.split) + ", \"coll_gather\": " +
                    std::to_string(best.gather) + ", \"xcd_run\": " + std::to_string(best.xcd) +
                    ", \"sort_items\": " + std::to_string(best.items) + ", \"fuse_bonds\": " + std::to_string(best.fuse) +
                    ", \"pass_waves\": " + std::to_string(best.waves) + "}, \"steps_per_candidate\": " + std::to_string(steps) + ", \"candidates\": [";
    for (size_t q = 0; q < tried.size(); q++) {
      char buf[240];
      std::snprintf(buf, sizeof(buf),
                    "%s{\"relax_split\": %d, \"coll_gather\": %d, \"xcd_run\": %d, \"sort_items\": %d, "
                    "\"fuse_bonds\": %d, \"pass_waves\": %d, \"ms_per_step\": %.4f}",
                    q ? ", " : "", tried[q].split, tried[q].gather, tried[q].xcd, tried[q].items, tried[q].fuse,
                    tried[q].waves, tried[q].ms);
      r += buf;
    }
    r += "]}";
    std::snprintf(report, report_cap, "%s", r.c_str());
  }
  return SC_OK;
}

sc_status sc_eval_primitive(int device, int fn, const float* x, float* y, size_t n, int on_device) {
  if (fn < 0 || fn > 11 || (n && (!x || !y))) return SC_ERR_INVALID;
  if (!on_device) {
    for (size_t i = 0; i < n; i++) y[i] = eval_primitive(fn, x[i]);
    return SC_OK;
  }
  if (hipSetDevice(device) != hipSuccess) return SC_ERR_DEVICE;
  float *dx = nullptr, *dy = nullptr;
  hipStream_t st_ = nullptr;
  if (hipStreamCreateWithFlags(&st_, hipStreamNonBlocking) != hipSuccess) return SC_ERR_DEVICE;
  if (hipMalloc(&dx, (n ? n : 1) * 4) != hipSuccess) { (void)hipStreamDestroy(st_); return SC_ERR_OOM; }
  if (hipMalloc(&dy, (n ? n : 1) * 4) != hipSuccess) { (void)hipFree(dx); (void)hipStreamDestroy(st_); return SC_ERR_OOM; }
  sc_status st = SC_OK;
  if (hipMemcpyAsync(dx, x, n * 4, hipMemcpyHostToDevice, st_) != hipSuccess) st = SC_ERR_DEVICE;
  if (st == SC_OK) {
    launch_eval_primitive(fn, dx, dy, n, st_);
    if (hipStreamSynchronize(st_) != hipSuccess) st = SC_ERR_DEVICE;
  }
  if (st == SC_OK && (hipMemcpyAsync(y, dy, n * 4, hipMemcpyDeviceToHost, st_) != hipSuccess || hipStreamSynchronize(st_) != hipSuccess)) st = SC_ERR_DEVICE;
  (void)hipFree(dx);
  (void)hipFree(dy);
  (void)hipStreamDestroy(st_);
  return st;
}

}  // extern "C"

// ============================================================================
// Serving: a batch with reserved capacity whose instances join and leave
// between steps (the scheduler is sc_server.cpp).
// ============================================================================

// Appends records [0, n) of array k at slots [offset, offset + n) of both
// buffer sets, with instance id `instance`.
static sc_status upload_slots(sc_sim* sim, int k, size_t offset, const sc_particle* src, size_t n, int instance) {
  Pop& p = sim->pop[k];
  if (n == 0) return SC_OK;
  if (offset + n > (size_t)p.cap) return fail(sim, SC_ERR_BUDGET, "no free slots");
  std::vector<float4> pos(n), vel(n), rot(n), avel(n), pn(n), an(n), dbg(n), m0, m1;
  std::vector<int4> hdr(n), st(n), bonds;
  std::vector<float2> m2;
  std::vector<int> ins(n, instance);
  if (p.metabolic) { m0.resize(n); m1.resize(n); m2.resize(n); }
  else bonds.assign(n * SC_MAX_BONDS, make_int4(0, 0, 0, 0));
  for (size_t i = 0; i < n; i++) {
    const sc_particle& r = src[i];
    pos[i] = make_float4(r.pos[0], r.pos[1], r.pos[2], r.radius);
    vel[i] = make_float4(r.vel[0], r.vel[1], r.vel[2], 0.0f);
    rot[i] = make_float4(r.rot[0], r.rot[1], r.rot[2], r.rot[3]);
    avel[i] = make_float4(r.ang_vel[0], r.ang_vel[1], r.ang_vel[2], r.ang_vel[3]);
    pn[i] = make_float4(r.pos_noise[0], r.pos_noise[1], r.pos_noise[2], 0.0f);
    an[i] = make_float4(r.ang_noise[0], r.ang_noise[1], r.ang_noise[2], r.ang_noise[3]);
    dbg[i] = make_float4(r.debug[0], r.debug[1], r.debug[2], r.debug[3]);
    hdr[i] = make_int4(r.id, r.type, r.flags, (int8_t)r.hydrophobic);
    if (r.n_bonds < 0 || r.n_bonds > SC_MAX_BONDS) return fail(sim, SC_ERR_INVALID, "n_bonds out of range");
    st[i] = make_int4((int8_t)r.state, r.last_state_change_step, r.n_bonds, r.scheduled_polymerization);
    if (p.metabolic) {
      m0[i] = make_float4(r.metabolites[0], r.metabolites[1], r.metabolites[2], r.metabolites[3]);
      m1[i] = make_float4(r.metabolites[4], r.metabolites[5], r.metabolites[6], r.metabolites[7]);
      m2[i] = make_float2(r.metabolites[8], r.metabolites[9]);
    } else {
      for (int b = 0; b < SC_MAX_BONDS; b++) bonds[(size_t)b * n + i] = make_int4(r.bonds[b][0], r.bonds[b][1], r.bonds[b][2], 0);
    }
  }
  // every piece packed into one pinned buffer, one copy to the device, then
  // device-side copies into both buffers of each field, one wait at the end
  struct Piece { Field f; size_t dst_off; const void* src; size_t bytes; };
  std::vector<Piece> pieces;
  auto put = [&](Field f, const void* data, size_t elem) { pieces.push_back({f, offset * elem, data, n * elem}); };
  put(F_POS, pos.data(), 16);
  put(F_VEL, vel.data(), 16);
  put(F_ROT, rot.data(), 16);
  put(F_AVEL, avel.data(), 16);
  put(F_PNOISE, pn.data(), 16);
  put(F_ANOISE, an.data(), 16);
  put(F_DBG, dbg.data(), 16);
  put(F_HDR, hdr.data(), 16);
  put(F_ST, st.data(), 16);
  put(F_INST, ins.data(), 4);
  if (p.metabolic) {
    put(F_MET0, m0.data(), 16);
    put(F_MET1, m1.data(), 16);
    put(F_MET2, m2.data(), 8);
  } else {
    for (int b = 0; b < SC_MAX_BONDS; b++)
      pieces.push_back({F_BONDS, ((size_t)b * p.cap + offset) * 16, &bonds[(size_t)b * n], n * 16});
  }
  size_t total = 0;
  std::vector<size_t> at(pieces.size());
  for (size_t q = 0; q < pieces.size(); q++) {
    at[q] = total;
    total += (pieces[q].bytes + 255) / 256 * 256;
  }
  if (total > sim->up_cap) {
    void* dev = nullptr;
    sc_status s = dalloc(sim, &dev, total);
    if (s != SC_OK) return s;
    void* host = nullptr;
    if (hipHostMalloc(&host, total) != hipSuccess) return fail(sim, SC_ERR_OOM, "pinned upload staging");
    if (sim->up_dev) {
      auto it = std::find(sim->allocs.begin(), sim->allocs.end(), sim->up_dev);
      if (it != sim->allocs.end()) sim->allocs.erase(it);
      SC_TRY(hipFree(sim->up_dev));
    }
    if (sim->up_host) SC_TRY(hipHostFree(sim->up_host));
    sim->up_dev = dev;
    sim->up_host = host;
    sim->up_cap = total;
  }
  for (size_t q = 0; q < pieces.size(); q++) std::memcpy((char*)sim->up_host + at[q], pieces[q].src, pieces[q].bytes);
  SC_TRY(hipMemcpyAsync(sim->up_dev, sim->up_host, total, hipMemcpyHostToDevice, sim->stream));
  for (size_t q = 0; q < pieces.size(); q++)
    for (int b = 0; b < 2; b++)
      SC_TRY(hipMemcpyAsync((char*)p.buf[pieces[q].f][b] + pieces[q].dst_off, (char*)sim->up_dev + at[q], pieces[q].bytes,
                            hipMemcpyDeviceToDevice, sim->stream));
  SC_TRY(hipStreamSynchronize(sim->stream));   // the staging is reused by the next upload
  return SC_OK;
}

// Places the live instances' cell keys: largest first, each aligned to its
// own size.  False when they do not fit below the sentinel.
static bool pack_keys(sc_sim* sim) {
  std::vector<int> order;
  for (int i = 0; i < (int)sim->table.size(); i++)
    if (sim->live[i] && !sim->table[i].evicted) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    return sim->table[a].bits != sim->table[b].bits ? sim->table[a].bits > sim->table[b].bits : a < b;
  });
  uint64_t at = 0;
  for (int i : order) {
    const uint64_t cells = (uint64_t)1 << (3 * sim->table[i].bits);
    at = (at + cells - 1) / cells * cells;
    if (at + cells > sim->g.sentinel) return false;
    sim->table[i].key_base = (uint32_t)at;
    at += cells;
  }
  return true;
}

namespace scb {

const char* sim_error(const sc_sim* sim) { return sim ? sim->err.c_str() : ""; }

double bytes_per_slot(bool metabolic) {
  size_t b = 0;
  for (int f = 0; f < F_COUNT; f++) b += 2 * field_bytes((Field)f, 1, metabolic);   // both buffer sets
  b += 3 * 4 + 3 * 4 + 4;                                  // keys, sort values, bonded list
  b += 2 * (kPairSlots * sizeof(uint2) + sizeof(int));      // neighbour lists
  b += 2 * sizeof(int);                                     // worklists
  b += sizeof(float4) + sizeof(uint32_t);                   // merged metabolic + plain array
  return (double)b;
}

bool server_compatible(const scs::HostSetup& templ, const scs::HostSetup& h, std::string& why) {
  return same_shared_constants(templ, h, why);
}

sc_status server_sim_create(const scs::HostSetup& templ, const ServerLimits& lim, int device, sc_sim** out,
                            std::string& err) {
  *out = nullptr;
  if (lim.cap_plain < 1 || lim.cap_metabolic < 1 || lim.max_instances < 1 || lim.max_cells < 64 ||
      lim.max_cells >= 0xfffffff0u) {
    err = "server limits out of range";
    return SC_ERR_INVALID;
  }
  sc_sim* sim = new sc_sim();
  sim->batch = true;
  sim->server = true;
  sim->inst.assign(1, templ);
  sim->host = templ;
  auto bail = [&](sc_status s) { err = sim->err; sc_destroy(sim); return s; };
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { sim->err = "no HIP device available"; return bail(SC_ERR_DEVICE); }
  if (device < 0 || device >= ndev) { sim->err = "device ordinal out of range"; return bail(SC_ERR_INVALID); }
  sim->device = device;
  if (hipSetDevice(device) != hipSuccess) { sim->err = "hipSetDevice failed"; return bail(SC_ERR_DEVICE); }
  if (hipStreamCreateWithFlags(&sim->stream, hipStreamNonBlocking) != hipSuccess) { sim->err = "stream"; return bail(SC_ERR_DEVICE); }
  if (hipStreamCreateWithFlags(&sim->side, hipStreamNonBlocking) != hipSuccess) { sim->err = "stream"; return bail(SC_ERR_DEVICE); }
  const sc_params& P = templ.P;
  GridP& g = sim->g;
  g = instance_grid(templ);
  g.bonds = 1;                 // requests may bring bonded particles
  g.sentinel = lim.max_cells;
  int key_bits = 1;
  while (key_bits < 32 && ((uint64_t)1 << key_bits) <= (uint64_t)lim.max_cells + 1) key_bits++;
  g.key_bits = key_bits;
  g.passes = (g.key_bits + kRadixBits - 1) / kRadixBits;
  g.drop_evicted = 1;
  GridP unused = g;
  unused.evicted = 1;
  sim->table.assign(lim.max_instances, unused);
  sim->live.assign(lim.max_instances, 0);
  Phys& ph = sim->ph;
  ph.movement_noise[0] = P.movement_noise_scale; ph.movement_noise[1] = P.metabolite_movement_noise_scale;
  ph.rotation_noise = P.rotation_noise_scale; ph.noise_model = P.noise_model;
  ph.se_translation = P.se_translation; ph.se_rotation = P.se_rotation; ph.velocity_decay = P.velocity_decay;
  ph.angular_velocity_decay = P.angular_velocity_decay; ph.seed_lo = P.seed_lo; ph.seed_hi = P.seed_hi;
  ph.coord_steps = P.noise_coordination_steps; ph.capacity_plain = lim.cap_plain;
  const size_t cells = (size_t)g.sentinel + 1;           // the cell starts reach the dropped tail
  sc_status s;
  if ((s = alloc_pop(sim, sim->pop[0], lim.cap_plain, false, cells)) != SC_OK) return bail(s);
  if ((s = alloc_pop(sim, sim->pop[1], lim.cap_metabolic, true, cells)) != SC_OK) return bail(s);
  if ((s = alloc_fuse(sim)) != SC_OK) return bail(s);
  const size_t mixed = (size_t)lim.cap_plain + (size_t)lim.cap_metabolic;
  if ((s = dalloc(sim, (void**)&sim->pop[1].mix_pos, mixed * sizeof(float4))) != SC_OK) return bail(s);
  if ((s = dalloc(sim, (void**)&sim->pop[1].mix_meta, mixed * sizeof(uint32_t))) != SC_OK) return bail(s);
  if ((s = dalloc(sim, (void**)&sim->ctr, sizeof(Counters))) != SC_OK) return bail(s);
  if ((s = dalloc(sim, (void**)&sim->d_table, lim.max_instances * sizeof(GridP))) != SC_OK) return bail(s);
  if ((s = upload(sim, sim->d_table, sim->table)) != SC_OK) return bail(s);
  if ((s = dalloc(sim, (void**)&sim->d_ictr, lim.max_instances * sizeof(InstCtr))) != SC_OK) return bail(s);
  g.table = sim->d_table;
  if ((s = upload_tables(sim, templ)) != SC_OK) return bail(s);
  sim->frame_bytes = (size_t)lim.cap_plain * SC_REDUCED_PARTICLE_BYTES + (size_t)lim.cap_metabolic * SC_REDUCED_METABOLIC_BYTES;
  if ((s = dalloc(sim, &sim->d_frame, sim->frame_bytes)) != SC_OK) return bail(s);
  sc_counts c0;
  std::memset(&c0, 0, sizeof(c0));
  if ((s = set_counts(sim, c0)) != SC_OK) return bail(s);
  *out = sim;
  return SC_OK;
}

sc_status server_admit(sc_sim* sim, int slot, const scs::HostSetup& h) {
  if (!sim || !sim->server || slot < 0 || slot >= (int)sim->table.size() || sim->live[slot])
    return fail(sim, SC_ERR_INVALID, "bad instance slot");
  SC_TRY(hipSetDevice(sim->device));
  Counters c;
  sc_status s = get_counters(sim, c);
  if (s != SC_OK) return s;
  if ((size_t)c.n_active[0] + h.plain.size() > (size_t)sim->pop[0].cap ||
      (size_t)c.n_active[1] + h.metabolic.size() > (size_t)sim->pop[1].cap)
    return fail(sim, SC_ERR_BUDGET, "no free slots");
  const GridP before = sim->table[slot];
  GridP gi = instance_grid(h);
  gi.sentinel = sim->g.sentinel; gi.key_bits = sim->g.key_bits; gi.passes = sim->g.passes;
  gi.step0 = c.step;
  gi.fresh = h.run_init ? 1 : 0;
  sim->table[slot] = gi;
  sim->live[slot] = 1;
  // Any failure from here on leaves the batch as it was: the slot unused,
  // the keys of the other instances where they were, the active counts
  // (once raised) back to the resident slots.  Particles already written
  // past those counts are outside every kernel's range.
  bool counts_raised = false;
  auto undo = [&](sc_status why) {
    const std::string msg = sim->err;
    sim->table[slot] = before;
    sim->live[slot] = 0;
    pack_keys(sim);
    if (counts_raised) (void)mem_copy(sim, sim->ctr->n_active, c.n_active, sizeof(c.n_active), hipMemcpyHostToDevice);
    (void)upload(sim, sim->d_table, sim->table);
    sim->err = msg;
    return why;
  };
  auto hip_ok = [&](hipError_t e, const char* what) {
    if (e == hipSuccess) return SC_OK;
    sim->err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? SC_ERR_OOM : SC_ERR_DEVICE;
  };
  if (!pack_keys(sim)) {
    (void)fail(sim, SC_ERR_BUDGET, "cells do not fit in the key space");
    return undo(SC_ERR_BUDGET);
  }
  if (sim->inst.size() < sim->table.size()) sim->inst.resize(sim->table.size(), sim->host);
  sim->inst[slot].P = h.P;                     // capacities and box for frames of this instance
  if ((s = upload_slots(sim, 0, (size_t)c.n_active[0], h.plain.data(), h.plain.size(), slot)) != SC_OK) return undo(s);
  if ((s = upload_slots(sim, 1, (size_t)c.n_active[1], h.metabolic.data(), h.metabolic.size(), slot)) != SC_OK)
    return undo(s);
  int na[2] = {c.n_active[0] + (int)h.plain.size(), c.n_active[1] + (int)h.metabolic.size()};
  counts_raised = true;
  if ((s = hip_ok(mem_copy(sim, sim->ctr->n_active, na, sizeof(na), hipMemcpyHostToDevice), "active counts")) != SC_OK)
    return undo(s);
  InstCtr ic;
  ic.next_particle_id = h.counts.next_particle_id;
  ic.n_plain = (int)h.plain.size();
  ic.cap_plain = h.P.num_particles;
  ic.n_metabolic = (int)h.metabolic.size();
  if ((s = hip_ok(mem_copy(sim, sim->d_ictr + slot, &ic, sizeof(ic), hipMemcpyHostToDevice), "instance counters")) !=
      SC_OK)
    return undo(s);
  if ((s = upload(sim, sim->d_table, sim->table)) != SC_OK) return undo(s);
  if (h.run_init) {
    // the init sequence acts on the fresh instance only; re-sorting the
    // others by unchanged keys leaves their order as it is
    if ((s = run_init(sim)) != SC_OK) return undo(s);
    sim->table[slot].fresh = 0;
    if ((s = upload(sim, sim->d_table, sim->table)) != SC_OK) return undo(s);
  }
  return SC_OK;
}

void server_set_launch(sc_sim* sim, size_t slots_plain, size_t slots_metabolic) {
  const size_t want[2] = {slots_plain, slots_metabolic};
  for (int k = 0; k < 2; k++) {
    size_t l = 256;
    while (l < want[k]) l *= 2;
    sim->pop[k].launch = (int)std::min(l, (size_t)std::max(sim->pop[k].cap, 1));
  }
}

sc_status server_evict(sc_sim* sim, int slot) {
  if (!sim || !sim->server || slot < 0 || slot >= (int)sim->table.size() || !sim->live[slot])
    return fail(sim, SC_ERR_INVALID, "bad instance slot");
  SC_TRY(hipSetDevice(sim->device));
  SC_TRY(hipStreamSynchronize(sim->stream));
  sim->table[slot].evicted = 1;
  sim->live[slot] = 0;
  pack_keys(sim);
  return upload(sim, sim->d_table, sim->table);
}

// ---- retiring without holding the batch -----------------------------------------
// Stage layout (device and host halves alike): a 256-byte header (int slot
// counts of both arrays at 0, the instance's InstCtr at 16), then per array
// its gathered fields, `cap` rows each, every field 256-byte aligned.
static size_t retire_layout(const size_t cap[2], size_t off[2][F_COUNT]) {
  size_t bytes = 256;
  for (int k = 0; k < 2; k++)
    for (int f = 0; f < F_COUNT; f++) {
      off[k][f] = bytes;
      bytes += (field_bytes((Field)f, std::max<size_t>(cap[k], 1), k == 1) + 255) / 256 * 256;
    }
  return bytes;
}

size_t server_retire_bytes(size_t cap_plain, size_t cap_metabolic) {
  const size_t cap[2] = {cap_plain, cap_metabolic};
  size_t off[2][F_COUNT];
  return retire_layout(cap, off);
}

sc_status server_stage_alloc(sc_sim* sim, size_t bytes, RetireStage& st) {
  SC_TRY(hipSetDevice(sim->device));
  st = RetireStage{};
  if (hipMalloc(&st.dev, bytes) != hipSuccess) return fail(sim, SC_ERR_OOM, "retire stage (device)");
  if (hipHostMalloc(&st.host, bytes) != hipSuccess) {
    (void)hipFree(st.dev);
    st.dev = nullptr;
    return fail(sim, SC_ERR_OOM, "retire stage (pinned host)");
  }
  hipEvent_t e = nullptr;
  SC_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  st.done = (void*)e;
  st.bytes = bytes;
  return SC_OK;
}

void server_stage_free(sc_sim* sim, RetireStage& st) {
  if (sim) (void)hipSetDevice(sim->device);
  if (st.done) {
    (void)hipEventSynchronize((hipEvent_t)st.done);   // a copy still in flight reads both halves
    (void)hipEventDestroy((hipEvent_t)st.done);
  }
  if (st.dev) (void)hipFree(st.dev);
  if (st.host) (void)hipHostFree(st.host);
  st = RetireStage{};
}

sc_status server_retire_begin(sc_sim* sim, int slot, const RetireStage& st, Retired& r) {
  if (!sim || !sim->server || slot < 0 || slot >= (int)sim->table.size() || !sim->live[slot])
    return fail(sim, SC_ERR_INVALID, "bad instance slot");
  size_t off[2][F_COUNT];
  const size_t need = retire_layout(r.cap, off);
  if (need > st.bytes) return fail(sim, SC_ERR_INVALID, "retire stage too small");
  SC_TRY(hipSetDevice(sim->device));
  Counters h;
  sc_status s = get_counters(sim, h);
  if (s != SC_OK) return s;
  r.step = h.step - sim->table[slot].step0;
  char* dev = (char*)st.dev;
  for (int k = 0; k < 2; k++) {
    const size_t n = (size_t)h.n_active[k];
    if (n == 0 || r.cap[k] == 0) {
      SC_TRY(hipMemsetAsync(dev + 4 * k, 0, 4, sim->stream));
      continue;
    }
    if ((s = islots_reserve(sim, n)) != SC_OK) return s;
    const size_t tiles = (n + kBlock - 1) / kBlock;
    int* tile_cnt = sim->islots;
    int* slots = tile_cnt + tiles;
    int* total = slots + n;
    launch_instance_slots(sim->pop[k], (int)n, slot, tile_cnt, slots, total, sim->stream);
    void* out[F_COUNT];
    for (int f = 0; f < F_COUNT; f++) out[f] = dev + off[k][f];
    launch_gather_dev(sim->pop[k], slots, total, (int)r.cap[k], out, sim->stream);
    // the count goes with the rows (the scratch is reused by the next search)
    SC_TRY(hipMemcpyAsync(dev + 4 * k, total, sizeof(int), hipMemcpyDeviceToDevice, sim->stream));
  }
  SC_TRY(hipMemcpyAsync(dev + 16, sim->d_ictr + slot, sizeof(InstCtr), hipMemcpyDeviceToDevice, sim->stream));
  SC_TRY(hipGetLastError());
  if (!sim->copy) SC_TRY(hipStreamCreateWithFlags(&sim->copy, hipStreamNonBlocking));
  if (!sim->retire_ev) SC_TRY(hipEventCreateWithFlags(&sim->retire_ev, hipEventDisableTiming));
  SC_TRY(hipEventRecord(sim->retire_ev, sim->stream));
  SC_TRY(hipStreamWaitEvent(sim->copy, sim->retire_ev, 0));
  SC_TRY(hipMemcpyAsync(st.host, st.dev, need, hipMemcpyDeviceToHost, sim->copy));
  SC_TRY(hipEventRecord((hipEvent_t)st.done, sim->copy));
  return SC_OK;
}

void server_drain(sc_sim* sim) {
  if (!sim) return;
  (void)hipSetDevice(sim->device);
  (void)hipStreamSynchronize(sim->stream);
  if (sim->copy) (void)hipStreamSynchronize(sim->copy);
}

bool server_retire_ready(const RetireStage& st) {
  return st.done && hipEventQuery((hipEvent_t)st.done) == hipSuccess;
}

void server_retire_counts(const RetireStage& st, const Retired& r, sc_counts* out) {
  InstCtr ic;
  std::memcpy(&ic, (const char*)st.host + 16, sizeof(ic));
  out->n_active = ic.n_plain;
  out->n_active_metabolic = ic.n_metabolic;
  out->next_particle_id = ic.next_particle_id;
  out->step = r.step;
}

void server_retire_decode(const RetireStage& st, const Retired& r, sc_particle* plain, size_t plain_cap,
                          sc_particle* metabolic, size_t metabolic_cap) {
  size_t off[2][F_COUNT];
  retire_layout(r.cap, off);
  const char* host = (const char*)st.host;
  for (int k = 0; k < 2; k++) {
    sc_particle* dst = k == 0 ? plain : metabolic;
    const size_t want = k == 0 ? plain_cap : metabolic_cap;
    if (!dst) continue;
    int m = 0;
    std::memcpy(&m, host + 4 * k, sizeof(int));
    const size_t n = std::min({(size_t)std::max(m, 0), want, r.cap[k]});
    const void* hs[F_COUNT];
    for (int f = 0; f < F_COUNT; f++) hs[f] = host + off[k][f];
    if (n) decode_fields(hs, r.cap[k], k == 1, n, dst);
    for (size_t i = n; i < want; i++) std::memset(&dst[i], 0, sizeof(sc_particle));
  }
}

}  // namespace scb
