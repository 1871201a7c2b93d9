// ORACLE -- test infrastructure only (parity checker and CPU baseline).
//
// Serial restatement of the per-step particle pipeline of spatial_cell_sim:
// one function per reference kernel, each looping over the slots exactly as
// the reference's one-thread-per-slot launch does, on AoS double buffers that
// are zeroed before every Jacobi pass (memory.cuh:154-157).  Its real type
// (ref_math.h) is float; built with ORACLE_FP64 it is double throughout the
// step, the bound on the float path's rounding.  Deliberate
// differences from the reference, each a definition of behaviour the
// reference leaves undefined or implementation-specific:
//   * RNG: Philox4x32-10 keyed by (seed) with counter (particle id, step,
//     array, block) instead of one cuRAND XORWOW state per slot
//     (main.cu:54-66, dynamics.cuh:176-207).
//   * Transcendentals: sc_fmath.h instead of CUDA's libdevice.
//   * Cell ranges: an inactive key writes nothing (grid.cuh:56-72 writes
//     gridRanges[-2]/[-1] in that case).
//   * Bond lists never grow past 8 entries (appends beyond are dropped).
//   * polymerize runs serially in slot order; slot and id of each new monomer
//     are assigned in that order; the end-of-chain branch without an active
//     polymerisation bond or without a found partner touches nothing outside
//     the controller (interactions.cuh:598,713-728 index particles[-1]).
//   * Interaction-table / type-table lookups outside the tables read a zero
//     row.
//   * Periodic boundary (extension): wrapped stencil over L/cell cells per
//     axis, neighbour positions shifted by +-L, positions wrapped into [0, L).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "ref_math.h"

// Pure Jacobi passes (each particle reads the current buffers and writes only
// its own record) may run in parallel over particles: with -fopenmp the
// results are the same, particle for particle, as the serial build's.
#ifdef _OPENMP
#define SC_ORACLE_PARALLEL _Pragma("omp parallel for schedule(dynamic, 64)")
#else
#define SC_ORACLE_PARALLEL
#endif

using namespace orc;

// Planted mistakes (test infrastructure): built with ORACLE_PLANTS, orc_plant(k)
// switches on mistake k in the function it names, so a test can show that
// the checks pinning that function fail on it.  In the ordinary builds
// PLANT(k) is the constant false and the code is the plain restatement.
#ifdef ORACLE_PLANTS
static int g_plant = 0;
#define PLANT(k) (g_plant == (k))
extern "C" void orc_plant(int k) { g_plant = k; }
#else
#define PLANT(k) false
#endif

namespace {

const uint32_t MAX_GRID_INDEX = 0xFFFFFFFFu;  // constants.cuh:12

struct Bond { int type, group, partnerId; };

// types.cuh:167-236 (the metabolites array is used by the metabolic array only)
struct Particle {
  int id, type, flags, state, lastStateChangeStep, hydrophobic, nActiveInteractions, scheduledPolymerization;
  real radius;
  float3 pos;
  float4 rot;
  float3 velocity;
  float3 posNoise;
  float4 angularVelocity;
  float4 angularNoise;
  float4 debugVector;
  Bond interactions[SC_MAX_BONDS];
  real metabolites[SC_NUM_METABOLITES];
};

Particle from_sc(const sc_particle& s) {
  Particle p;
  std::memset(&p, 0, sizeof(p));
  p.id = s.id; p.type = s.type; p.flags = s.flags; p.state = (int8_t)s.state;
  p.lastStateChangeStep = s.last_state_change_step; p.hydrophobic = (int8_t)s.hydrophobic;
  p.nActiveInteractions = s.n_bonds; p.scheduledPolymerization = s.scheduled_polymerization;
  p.radius = s.radius;
  p.pos = mk3(s.pos[0], s.pos[1], s.pos[2]);
  p.rot = mk4(s.rot[0], s.rot[1], s.rot[2], s.rot[3]);
  p.velocity = mk3(s.vel[0], s.vel[1], s.vel[2]);
  p.posNoise = mk3(s.pos_noise[0], s.pos_noise[1], s.pos_noise[2]);
  p.angularVelocity = mk4(s.ang_vel[0], s.ang_vel[1], s.ang_vel[2], s.ang_vel[3]);
  p.angularNoise = mk4(s.ang_noise[0], s.ang_noise[1], s.ang_noise[2], s.ang_noise[3]);
  p.debugVector = mk4(s.debug[0], s.debug[1], s.debug[2], s.debug[3]);
  for (int k = 0; k < SC_MAX_BONDS; k++) {
    p.interactions[k].type = s.bonds[k][0];
    p.interactions[k].group = s.bonds[k][1];
    p.interactions[k].partnerId = s.bonds[k][2];
  }
  for (int k = 0; k < SC_NUM_METABOLITES; k++) p.metabolites[k] = s.metabolites[k];
  return p;
}

void to_sc(const Particle& p, sc_particle& s, bool metabolic) {
  std::memset(&s, 0, sizeof(s));
  s.id = p.id; s.type = p.type; s.flags = p.flags; s.state = p.state;
  s.last_state_change_step = p.lastStateChangeStep; s.hydrophobic = p.hydrophobic;
  s.n_bonds = p.nActiveInteractions; s.scheduled_polymerization = p.scheduledPolymerization;
  s.radius = p.radius;
  s.pos[0] = p.pos.x; s.pos[1] = p.pos.y; s.pos[2] = p.pos.z;
  s.rot[0] = p.rot.x; s.rot[1] = p.rot.y; s.rot[2] = p.rot.z; s.rot[3] = p.rot.w;
  s.vel[0] = p.velocity.x; s.vel[1] = p.velocity.y; s.vel[2] = p.velocity.z;
  s.pos_noise[0] = p.posNoise.x; s.pos_noise[1] = p.posNoise.y; s.pos_noise[2] = p.posNoise.z;
  s.ang_vel[0] = p.angularVelocity.x; s.ang_vel[1] = p.angularVelocity.y;
  s.ang_vel[2] = p.angularVelocity.z; s.ang_vel[3] = p.angularVelocity.w;
  s.ang_noise[0] = p.angularNoise.x; s.ang_noise[1] = p.angularNoise.y;
  s.ang_noise[2] = p.angularNoise.z; s.ang_noise[3] = p.angularNoise.w;
  s.debug[0] = p.debugVector.x; s.debug[1] = p.debugVector.y;
  s.debug[2] = p.debugVector.z; s.debug[3] = p.debugVector.w;
  // entries beyond nActiveInteractions are not state (the reference leaves
  // stale copies there after a removal); the snapshot reports them as zero
  for (int k = 0; k < SC_MAX_BONDS && k < p.nActiveInteractions; k++) {
    s.bonds[k][0] = p.interactions[k].type;
    s.bonds[k][1] = p.interactions[k].group;
    s.bonds[k][2] = p.interactions[k].partnerId;
  }
  if (metabolic)
    for (int k = 0; k < SC_NUM_METABOLITES; k++) s.metabolites[k] = p.metabolites[k];
}

inline bool active(const Particle& p) { return (p.flags & SC_FLAG_ACTIVE) != 0; }

// interactions.cuh:186-189
inline bool participantOrder(const Particle& p, const Particle& tp) {
  return p.type < tp.type || (p.type == tp.type && p.id < tp.id);
}
// interactions.cuh:191-197
inline float4 targetRelativeOrientation(const Particle& p, const Particle& tp, float4 q) {
  return participantOrder(p, tp) ? qmul(tp.rot, inverse(q)) : qmul(tp.rot, q);
}
// interactions.cuh:199-213
inline float3 relaxedRelativePositionT0(const Particle& p, const Particle& tp, float4 tro, float4 trpr, real dist) {
  float3 r = participantOrder(p, tp)
                 ? -transform_vector(VEC_UP(), qmul(tp.rot, qmul(inverse(tro), trpr)))
                 : transform_vector(VEC_UP(), qmul(tp.rot, trpr));
  r *= dist;
  r += tp.pos;
  return r;
}

bool append_bond(Particle& p, int type, int group, int partner) {
  if (p.nActiveInteractions >= SC_MAX_BONDS) return false;
  p.interactions[p.nActiveInteractions].type = type;
  p.interactions[p.nActiveInteractions].group = group;
  p.interactions[p.nActiveInteractions].partnerId = partner;
  p.nActiveInteractions++;
  return true;
}

}  // namespace

struct orc_sim {
  sc_params P;
  int nGridCells, nGridCellsBits, nPer;
  std::vector<sc_type_info> types;
  std::vector<sc_interaction_info> flat, partner;
  std::vector<int32_t> pidx;
  std::vector<sc_reaction> reactions;
  std::vector<Particle> cur[2], nxt[2];
  // site bonds (extension): per plain slot the chosen partner of each of up
  // to kSiteChoices groups, {row id, partner slot} pairs, -1 when none
  std::vector<int32_t> siteChoice;
  std::vector<uint32_t> sortedKeys[2];
  std::vector<int32_t> perm[2];
  std::vector<uint32_t> ranges[2];
  int nSorted[2];
  int nActive[2];
  int nextParticleId;
  int step;

  int capacity(int pop) const { return pop == 0 ? P.num_particles : P.num_metabolic_particles; }

  sc_interaction_info info(int id) const {
    if (id >= 0 && id < (int)flat.size()) return flat[id];
    sc_interaction_info z; std::memset(&z, 0, sizeof(z)); return z;
  }

  // grid.cuh:13-17 (truncate, then clamp to the last cell)
  uint32_t gridIdx(real c) const {
    uint32_t g = (c >= 0.0f) ? (uint32_t)(c / P.grid_cell_size) : 0u;
    uint32_t lim = (P.boundary == SC_BOUNDARY_PERIODIC) ? (uint32_t)(nPer - 1) : (uint32_t)(nGridCells - 1);
    return g < lim ? g : lim;
  }
  // grid.cuh:19-25
  uint32_t makeIdx(uint32_t gx, uint32_t gy, uint32_t gz) const {
    return (gx << (2 * nGridCellsBits)) | (gy << nGridCellsBits) | gz;
  }

  // Visits the 27-cell stencil in the reference order (gx, gy, gz ascending,
  // clipped, dynamics.cuh:56-62); periodic mode visits offsets -1, 0, +1 with
  // wrap-around and reports the image shift.
  template <class F>
  void forStencil(float3 pos, F&& f) const {
    int cg[3] = {(int)gridIdx(pos.x), (int)gridIdx(pos.y), (int)gridIdx(pos.z)};
    if (P.boundary != SC_BOUNDARY_PERIODIC) {
      for (int gx = std::max(cg[0] - 1, 0); gx <= std::min(cg[0] + 1, nGridCells - 1); gx++)
        for (int gy = std::max(cg[1] - 1, 0); gy <= std::min(cg[1] + 1, nGridCells - 1); gy++)
          for (int gz = std::max(cg[2] - 1, 0); gz <= std::min(cg[2] + 1, nGridCells - 1); gz++)
            f(makeIdx(gx, gy, gz), 0.0f, 0.0f, 0.0f);
      return;
    }
    real L = P.sim_size;
    for (int dx = -1; dx <= 1; dx++) {
      int gx = cg[0] + dx; real sx = 0.0f;
      if (gx < 0) { gx += nPer; sx = -L; } else if (gx >= nPer) { gx -= nPer; sx = L; }
      for (int dy = -1; dy <= 1; dy++) {
        int gy = cg[1] + dy; real sy = 0.0f;
        if (gy < 0) { gy += nPer; sy = -L; } else if (gy >= nPer) { gy -= nPer; sy = L; }
        for (int dz = -1; dz <= 1; dz++) {
          int gz = cg[2] + dz; real sz = 0.0f;
          if (gz < 0) { gz += nPer; sz = -L; } else if (gz >= nPer) { gz -= nPer; sz = L; }
          f(makeIdx(gx, gy, gz), sx, sy, sz);
        }
      }
    }
  }

  // Neighbour enumeration: every slot j of the stencil cells of `pop`, with the
  // neighbour's position shifted to the image nearest the stencil.
  template <class F>
  void forNeighbors(int pop, float3 pos, F&& f) const {
    const std::vector<uint32_t>& R = ranges[pop];
    const std::vector<Particle>& C = cur[pop];
    forStencil(pos, [&](uint32_t cell, real sx, real sy, real sz) {
      uint32_t s = R[2 * (size_t)cell], e = R[2 * (size_t)cell + 1];
      for (uint32_t j = s; j < e; j++) {
        const Particle& tp0 = C[j];
        if (sx != 0.0f || sy != 0.0f || sz != 0.0f) {
          Particle tp = tp0;
          tp.pos = mk3(tp.pos.x + sx, tp.pos.y + sy, tp.pos.z + sz);
          f((int)j, tp);
        } else {
          f((int)j, tp0);
        }
      }
    });
  }

  // Cell-by-cell variant for kernels that interleave two arrays per cell.
  template <class F>
  void forCells(float3 pos, F&& f) const { forStencil(pos, f); }

  static Particle shifted(const Particle& tp, real sx, real sy, real sz) {
    if (sx == 0.0f && sy == 0.0f && sz == 0.0f) return tp;
    Particle t = tp;
    t.pos = mk3(t.pos.x + sx, t.pos.y + sy, t.pos.z + sz);
    return t;
  }

  bool outside(float3 d, real lim) const {
    return scm::fabs_(d.x) > lim || scm::fabs_(d.y) > lim || scm::fabs_(d.z) > lim;
  }

  void beginJacobi(int pop) {
    // memset of the next buffer (memory.cuh:154-157): slots never written stay zero
    std::vector<Particle>& N = nxt[pop];
    std::memset(N.data() + nActive[pop], 0, sizeof(Particle) * (size_t)(N.size() - nActive[pop]));
  }
  void endJacobi(int pop) { cur[pop].swap(nxt[pop]); }

  // ---- grid.cuh:34-44, 76-139 -----------------------------------------------
  void updateGridAndSort(int pop) {
    int n = capacity(pop);
    std::vector<Particle>& C = cur[pop];
    std::vector<uint32_t> keys(n);
    for (int i = 0; i < n; i++)
      keys[i] = active(C[i]) ? makeIdx(gridIdx(C[i].pos.x), gridIdx(C[i].pos.y), gridIdx(C[i].pos.z))
                             : MAX_GRID_INDEX;
    std::vector<int32_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    std::vector<Particle>& N = nxt[pop];
    sortedKeys[pop].resize(n);
    for (int i = 0; i < n; i++) { N[i] = C[order[i]]; sortedKeys[pop][i] = keys[order[i]]; }
    C.swap(N);
    perm[pop] = order;
    nSorted[pop] = nActive[pop];
    // grid.cuh:46-73 restated as first/last slot of each run of equal keys
    std::vector<uint32_t>& R = ranges[pop];
    std::fill(R.begin(), R.end(), 0u);
    const std::vector<uint32_t>& K = sortedKeys[pop];
    for (int i = 0; i < n; i++) {
      uint32_t k = K[i];
      if (k == MAX_GRID_INDEX) break;
      if (i == 0 || K[i - 1] != k) R[2 * (size_t)k] = (uint32_t)i;
      if (i == n - 1 || K[i + 1] != k) R[2 * (size_t)k + 1] = (uint32_t)(i + 1);
    }
  }

  // ---- dynamics.cuh:19-154 (net effect: debugVector.xyz = 0) -------------------
  void moveKernel(int pop) {
    beginJacobi(pop);
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[pop]; idx++) {
      Particle p = cur[pop][idx];
      if (active(p) && !PLANT(31)) { p.debugVector.x = 0; p.debugVector.y = 0; p.debugVector.z = 0; }
      if (PLANT(32)) p.debugVector.w = 0;      // planted: the whole vector reset
      if (PLANT(33) && !active(p)) p.debugVector = mk4(0, 0, 0, p.debugVector.w);   // planted: inactive reset too
      nxt[pop][idx] = p;
    }
    endJacobi(pop);
  }

  // ---- dynamics.cuh:156-215 -------------------------------------------------
  void brownian(int pop, real movementNoiseScale) {
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[pop]; idx++) {
      Particle& p = cur[pop][idx];
      if (!active(p)) continue;
      scm::u32x4 c0 = {(uint32_t)p.id, (uint32_t)step, (uint32_t)pop, 0u};
      scm::u32x4 c1 = {(uint32_t)p.id, (uint32_t)step, (uint32_t)pop, 1u};
      scm::u32x4 r0 = scm::philox4x32_10(c0, P.seed_lo, P.seed_hi);
      scm::u32x4 r1 = scm::philox4x32_10(c1, P.seed_lo, P.seed_hi);
      float4 dir = random_rotation(scm::u01_(r0.x), scm::u01_(r0.y), scm::u01_(r0.z));
      real nrm = scm::box_muller_(scm::u01_(r0.w), scm::u01_(r1.x));
      float4 rr = random_rotation(scm::u01_(r1.y), scm::u01_(r1.z), scm::u01_(r1.w));
      if (P.noise_model == SC_NOISE_STOKES_EINSTEIN) {
        // extension: Stokes-Einstein scales of a sphere of this radius (sc_params)
        const real r = p.radius;
        real mag = scm::fabs_(nrm) * scm::sqrt_((PLANT(41) ? P.se_translation / 3.0f : P.se_translation) / r);
        p.posNoise = transform_vector(VEC_UP(), dir) * mag;
        p.angularNoise = slerp(p.angularNoise, rr, scm::sqrt_(P.se_rotation / (PLANT(42) ? r : r * r * r)));
        continue;
      }
      real mag = (real)((double)(scm::fabs_(nrm) * movementNoiseScale) /
                          scm::dmax_((double)p.radius / 2.5, 1.0));
      p.posNoise = transform_vector(VEC_UP(), dir) * mag;
      p.angularNoise = slerp(p.angularNoise, rr,
                             (real)((double)P.rotation_noise_scale / ((double)p.radius / 2.5)));
    }
  }

  // ---- dynamics.cuh:217-242 -------------------------------------------------
  void applyNoise(int pop) {
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[pop]; idx++) {
      Particle& p = cur[pop][idx];
      if (!active(p)) continue;
      p.velocity += p.posNoise;
      p.posNoise = VEC_ZERO();
      p.angularVelocity = qmul(p.angularNoise, p.angularVelocity);
      p.angularNoise = QUAT_ID();
    }
  }

  real wrap(real x) const {
    real L = P.sim_size;
    real y = x - L * scm::floor_(x / L);
    if (PLANT(21) ? y > L : y >= L) y = y - L;    // planted 21: a position landing on L stays L
    if (y < 0.0f) y = 0.0f;
    return y;
  }

  // ---- dynamics.cuh:244-282 -------------------------------------------------
  void applyVelocities(int pop, real stepFraction) {
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[pop]; idx++) {
      Particle& p = cur[pop][idx];
      if (!active(p)) continue;
      float3 np = p.pos + p.velocity * stepFraction;
      if (P.boundary == SC_BOUNDARY_PERIODIC)
        p.pos = mk3(wrap(np.x), wrap(np.y), PLANT(22) ? clampf(np.z, 0, P.sim_size) : wrap(np.z));
      else
        p.pos = clamp3(np, 0, P.sim_size);
      p.velocity *= P.velocity_decay;
      p.rot = qmul(slerp(QUAT_ID(), p.angularVelocity, stepFraction), p.rot);
      p.rot /= length(p.rot);
      p.angularVelocity = slerp(p.angularVelocity, QUAT_ID(), 1.0f - P.angular_velocity_decay);
    }
  }

  // ---- dynamics.cuh:288-402 -------------------------------------------------
  void coordinateNoise(real coordinationFraction) {
    beginJacobi(0);
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[0]; idx++) {
      Particle p = cur[0][idx];
      if (!active(p) || p.nActiveInteractions < 1) { nxt[0][idx] = p; continue; }
      float3 dPosNoise = VEC_ZERO();
      float4 dAngularNoise = QUAT_ID();
      real countedInteractionWeight = 0.0f;
      forNeighbors(0, p.pos, [&](int j, const Particle& tp) {
        if (j == idx || !active(tp)) return;
        float3 delta = tp.pos - p.pos;
        if (outside(delta, P.max_interaction_distance)) return;
        for (int k = 0; k < p.nActiveInteractions; k++) {
          Bond b = p.interactions[k];
          if (b.type < 10 && coordinationFraction > 0.25f) continue;
          if (b.partnerId == tp.id) {
            real pPosNoiseMag = length(p.posNoise);
            real w = 1.0f;
            dPosNoise += (tp.posNoise - p.posNoise) * 0.9f * tp.radius / (p.radius + tp.radius) * w;
            dPosNoise *= scm::fmax_(length(dPosNoise), pPosNoiseMag) / length(dPosNoise);
            dAngularNoise = slerp(dAngularNoise, qmul(inverse(p.angularNoise), tp.angularNoise),
                                  0.9f * tp.radius / (p.radius + tp.radius) * w);
            countedInteractionWeight += w;
          }
        }
      });
      p.posNoise += dPosNoise / scm::fmax_(countedInteractionWeight, 1.0f);
      p.angularNoise = qmul(slerp(QUAT_ID(), dAngularNoise, 1.0f / scm::fmax_(countedInteractionWeight, 1.0f)),
                            p.angularNoise);
      nxt[0][idx] = p;
    }
    endJacobi(0);
  }

  // ---- dynamics.cuh:405-822 -------------------------------------------------
  void relax() {
    beginJacobi(0);
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[0]; idx++) {
      Particle p = cur[0][idx];
      if (!active(p)) { nxt[0][idx] = p; continue; }
      float3 dVelocity = VEC_ZERO();
      float4 dAngularVelocity = QUAT_ID();
      float3 phobicAttractionPos = VEC_ZERO();
      real phobicAttractionPowerSum = 0.0f;
      real countedInteractionWeight = 0.0f;
      float3 up = transform_vector(VEC_UP(), p.rot);
      float3 nextUp = transform_vector(up, p.angularVelocity);
      forNeighbors(0, p.pos, [&](int j, const Particle& tp) {
        if (j == idx || !active(tp)) return;
        float3 delta = tp.pos - p.pos;
        if (outside(delta, P.max_interaction_distance)) return;
        float3 normalizedDelta = normalize(delta);
        real dist = norm(delta);
        float3 nextDelta = tp.pos + tp.velocity - (p.pos + p.velocity);
        float3 normalizedNextDelta = normalize(nextDelta);
        real interactionDistance = p.radius + tp.radius;
        float3 tup = transform_vector(VEC_UP(), tp.rot);
        float3 nextTup = transform_vector(tup, p.angularVelocity);   // p's angular velocity (reference quirk)

        if ((p.hydrophobic > 0 || tp.hydrophobic > 0) &&
            (double)dist <= (double)(p.radius + tp.radius) + 0.8 * (double)(p.radius + tp.radius) * 0.5) {
          if (p.hydrophobic == SC_POLAR && tp.hydrophobic == SC_POLAR) {
            float3 pPhobicDir = -nextUp;
            float3 tpPhobicDir = -nextTup;
            float3 pPhobicDirVert = pPhobicDir * p.radius + p.pos + p.velocity;
            float3 tpPhobicDirVert = tpPhobicDir * tp.radius + tp.pos + tp.velocity;
            const real powerCap = 10.0f;
            real attractionPower = (real)(scm::dmin_(scm::dmax_((double)(length(tpPhobicDirVert - pPhobicDirVert) /
                                                                            (p.radius + tp.radius)), 1e-6),
                                                       2.0) / 2.0);
            const real distanceRelaxationSpeed = 0.05f;
            real w = 1.0f;
            countedInteractionWeight += w;
            dVelocity += (tpPhobicDirVert - pPhobicDirVert) * distanceRelaxationSpeed * attractionPower * w;
            real phobicAttractionPower =
                scm::pow_(scm::fmin_(1.0f / scm::fmax_(length(tpPhobicDirVert - pPhobicDirVert) / (p.radius + tp.radius), 1e-6f),
                                     powerCap) / powerCap,
                          0.2f);
            phobicAttractionPowerSum += phobicAttractionPower;
            phobicAttractionPos += tpPhobicDirVert * phobicAttractionPower;
          }
        }

        bool interactionPartners = false;
        for (int k = 0; k < p.nActiveInteractions; k++) {
          Bond b = p.interactions[k];
          if (b.partnerId != tp.id) continue;
          interactionPartners = true;
          if (b.type >= 10) {
            sc_interaction_info pii = info(b.type);
            float4 ro = mk4(pii.relative_orientation[0], pii.relative_orientation[1],
                            pii.relative_orientation[2], pii.relative_orientation[3]);
            float3 rp = mk3(pii.relative_position[0], pii.relative_position[1], pii.relative_position[2]);
            const real relativeOrientationRelaxationSpeed = 0.9f;
            float4 rel = qmul(inverse(qmul(p.angularVelocity, p.rot)),
                              qmul(tp.angularVelocity, targetRelativeOrientation(p, tp, ro)));
            float4 other = qmul(inverse(qmul(tp.angularVelocity, tp.rot)),
                                qmul(p.angularVelocity, targetRelativeOrientation(tp, p, ro)));
            dAngularVelocity = slerp(dAngularVelocity, qmul(slerp(rel, other, 0.5f), rel),
                                     relativeOrientationRelaxationSpeed * tp.radius / (p.radius + tp.radius));
            const real relativePositionRelaxationSpeed = 0.9f;
            float3 relaxedRelativePosition =
                participantOrder(p, tp) ? transform_vector(-rp, qmul(inverse(ro), qmul(tp.angularVelocity, tp.rot)))
                                        : transform_vector(rp, qmul(tp.angularVelocity, tp.rot));
            relaxedRelativePosition += tp.pos;
            float3 otherRelaxedRelativePosition =
                participantOrder(tp, p) ? transform_vector(-rp, qmul(inverse(ro), qmul(p.angularVelocity, p.rot)))
                                        : transform_vector(rp, qmul(p.angularVelocity, p.rot));
            otherRelaxedRelativePosition += p.pos;
            real w = 1.0f;
            countedInteractionWeight += w;
            float3 pDV = (relaxedRelativePosition + tp.velocity - (p.pos + p.velocity)) *
                         relativePositionRelaxationSpeed * tp.radius / (p.radius + tp.radius) * w;
            float3 tpDV = (otherRelaxedRelativePosition + p.velocity - (tp.pos + tp.velocity)) *
                          relativePositionRelaxationSpeed * p.radius / (p.radius + tp.radius) * w;
            dVelocity += (pDV - (pDV + tpDV) / 2.0f) * relativePositionRelaxationSpeed * p.radius /
                         (p.radius + tp.radius) * w;
          }
          if (b.type == 0) {
            float4 trod = qmul(quaternion(VEC_UP(), -PI_F / 6), quaternion(VEC_RIGHT(), PI_F / 6));
            const real relativeOrientationRelaxationSpeed = 0.2f;
            p.rot = slerp(p.rot, targetRelativeOrientation(p, tp, trod), relativeOrientationRelaxationSpeed);
            float4 trpr = quaternion(VEC_FORWARD(), -PI_F / 6);
            const real relativePositionRelaxationSpeed = 0.2f;
            float3 rr = relaxedRelativePositionT0(p, tp, trod, trpr, interactionDistance);
            p.debugVector.x = (real)((double)p.debugVector.x + (double)(rr.x - p.pos.x) * (1.0 - (double)relativePositionRelaxationSpeed));
            p.debugVector.y = (real)((double)p.debugVector.y + (double)(rr.y - p.pos.y) * (1.0 - (double)relativePositionRelaxationSpeed));
            p.debugVector.z = (real)((double)p.debugVector.z + (double)(rr.z - p.pos.z) * (1.0 - (double)relativePositionRelaxationSpeed));
          } else if (b.type == SC_INTERACTION_TYPE_CHAIN) {
            const real relativeOrientationRelaxationSpeed = 0.9f;
            real sgnF = (b.group == SC_GROUP_FORWARD) ? -1.0f : 1.0f;
            dAngularVelocity = slerp(dAngularVelocity, quaternionFromTo(nextUp, sgnF * normalizedNextDelta),
                                     relativeOrientationRelaxationSpeed * tp.radius / (p.radius + tp.radius));
            real flex = PI_F / 8;
            const real relativePositionRelaxationSpeed = 0.9f;
            float3 relUpVec = ((b.group == SC_GROUP_FORWARD) ? 1.0f : -1.0f) * nextTup;
            float3 axis = safeCross(relUpVec, -normalizedNextDelta);
            real cur_a = angle(relUpVec, -normalizedNextDelta);
            float3 relaxed = transform_vector(relUpVec, quaternion(axis, scm::fmin_(scm::fmax_(cur_a, -flex), flex)));
            relaxed *= interactionDistance;
            relaxed += tp.pos;
            float3 otherRelUpVec = sgnF * nextUp;
            float3 otherAxis = safeCross(otherRelUpVec, normalizedNextDelta);
            real other_a = angle(otherRelUpVec, normalizedNextDelta);
            float3 otherRelaxed = transform_vector(otherRelUpVec, quaternion(otherAxis, scm::fmin_(scm::fmax_(other_a, -flex), flex)));
            otherRelaxed *= interactionDistance;
            otherRelaxed += p.pos;
            real w = 1.0f;
            countedInteractionWeight += w;
            float3 pDV = (relaxed + tp.velocity - (p.pos + p.velocity)) * relativePositionRelaxationSpeed *
                         tp.radius / (p.radius + tp.radius) * w;
            float3 tpDV = (otherRelaxed + p.velocity - (tp.pos + tp.velocity)) * relativePositionRelaxationSpeed *
                          p.radius / (p.radius + tp.radius) * w;
            dVelocity += (pDV - (pDV + tpDV) / 2.0f) * relativePositionRelaxationSpeed * p.radius /
                         (p.radius + tp.radius) * w;
            p.debugVector.x = (real)((double)p.debugVector.x + (double)(relaxed.x - p.pos.x) * (1.0 - (double)relativePositionRelaxationSpeed));
            p.debugVector.y = (real)((double)p.debugVector.y + (double)(relaxed.y - p.pos.y) * (1.0 - (double)relativePositionRelaxationSpeed));
            p.debugVector.z = (real)((double)p.debugVector.z + (double)(relaxed.z - p.pos.z) * (1.0 - (double)relativePositionRelaxationSpeed));
          }
        }

        if (!interactionPartners) {
          real collisionDist = p.radius + tp.radius;
          if (dist < collisionDist) {
            real w = 1.0f;
            countedInteractionWeight += w;
            dVelocity += (tp.pos - normalizedDelta * collisionDist + tp.velocity - (p.pos + p.velocity)) * 0.9f *
                         tp.radius / (p.radius + tp.radius) * w;
          }
        }
      });

      p.velocity += dVelocity / scm::fmax_(countedInteractionWeight, 1.0f);
      p.angularVelocity = qmul(slerp(QUAT_ID(), dAngularVelocity, 1.0f / scm::fmax_(countedInteractionWeight, 1.0f)),
                               p.angularVelocity);
      if (phobicAttractionPowerSum > 0.0f) {
        float3 dir = normalize(phobicAttractionPos / phobicAttractionPowerSum - (p.pos + p.velocity));
        p.angularVelocity = slerp(p.angularVelocity, quaternionFromTo(-up, dir), 0.2f);
      }
      nxt[0][idx] = p;
    }
    endJacobi(0);
  }

  // ---- metabolites.cuh:140-264 ----------------------------------------------
  void relaxMetabolicParticles() {
    beginJacobi(1);
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[1]; idx++) {
      Particle p = cur[1][idx];
      if (!active(p)) { nxt[1][idx] = p; continue; }
      float3 dVelocity = VEC_ZERO();
      real countedInteractionWeight = 0.0f;
      forCells(p.pos, [&](uint32_t cell, real sx, real sy, real sz) {
        uint32_t ms = ranges[1][2 * (size_t)cell], me = ranges[1][2 * (size_t)cell + 1];
        for (uint32_t j = ms; j < me; j++) {
          if ((int)j == idx) continue;
          Particle tp = shifted(cur[1][j], sx, sy, sz);
          if (!active(tp)) continue;
          float3 delta = tp.pos - p.pos;
          if (outside(delta, P.max_interaction_distance)) continue;
          float3 nd = normalize(delta);
          real dist = norm(delta);
          real collisionDist = (real)((double)(p.radius + tp.radius) + 0.8 * (double)p.radius);
          if (dist <= collisionDist) {
            real w = 1.0f;
            countedInteractionWeight += w;
            dVelocity += (tp.pos - nd * collisionDist + tp.velocity - (p.pos + p.velocity)) * 0.9f * w;
          }
        }
        uint32_t ps = ranges[0][2 * (size_t)cell], pe = ranges[0][2 * (size_t)cell + 1];
        for (uint32_t j = ps; j < pe; j++) {
          Particle tp = shifted(cur[0][j], sx, sy, sz);
          if (!active(tp) || tp.type != SC_TYPE_LIPID) continue;
          float3 delta = tp.pos - p.pos;
          if (outside(delta, P.max_interaction_distance)) continue;
          float3 nd = normalize(delta);
          real dist = norm(delta);
          real collisionDist = (real)((double)(p.radius + tp.radius) + 0.8 * (double)p.radius);
          if (dist <= collisionDist) {
            real w = 2.0f;
            countedInteractionWeight += w;
            dVelocity += (tp.pos - nd * collisionDist + tp.velocity - (p.pos + p.velocity)) * 0.9f * w;
          }
        }
      });
      p.velocity += dVelocity / scm::fmax_(countedInteractionWeight, 1.0f);
      nxt[1][idx] = p;
    }
    endJacobi(1);
  }

  // ---- metabolites.cuh:266-350 ----------------------------------------------
  void relaxMetabolicParticlePartners() {
    beginJacobi(0);
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[0]; idx++) {
      Particle p = cur[0][idx];
      if (!active(p) || p.type != SC_TYPE_LIPID) { nxt[0][idx] = p; continue; }
      float3 dVelocity = VEC_ZERO();
      real countedInteractionWeight = 0.0f;
      forNeighbors(1, p.pos, [&](int j, const Particle& tp) {
        (void)j;
        if (!active(tp)) return;
        float3 delta = tp.pos - p.pos;
        if (outside(delta, P.max_interaction_distance)) return;
        float3 nd = normalize(delta);
        real dist = norm(delta);
        real collisionDist = (real)((double)(p.radius + tp.radius) + 0.8 * (double)p.radius);
        if (dist <= collisionDist) {
          real w = 2.0f;
          countedInteractionWeight += w;
          dVelocity += (tp.pos - nd * collisionDist + tp.velocity - (p.pos + p.velocity)) * 0.9f * w;
        }
      });
      p.velocity += dVelocity / scm::fmax_(countedInteractionWeight, 1.0f);
      nxt[0][idx] = p;
    }
    endJacobi(0);
  }

  // ---- metabolites.cuh:352-431 ----------------------------------------------
  void diffuseMetabolites() {
    beginJacobi(1);
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[1]; idx++) {
      Particle p = cur[1][idx];
      if (!active(p)) { nxt[1][idx] = p; continue; }
      real delta_m[SC_NUM_METABOLITES] = {0};
      forNeighbors(1, p.pos, [&](int j, const Particle& tp) {
        if (j == idx || !active(tp)) return;
        float3 delta = tp.pos - p.pos;
        if (outside(delta, P.max_diffusion_distance)) return;
        real dist = norm(delta);
        if (dist <= P.max_diffusion_distance) {
          for (int k = 0; k < SC_NUM_METABOLITES; k++) {
            real diff = tp.metabolites[k] - p.metabolites[k];
            if ((double)scm::fabs_(diff) > 0.001) delta_m[k] += diff * 0.1f;
          }
        }
      });
      for (int k = 0; k < SC_NUM_METABOLITES; k++) p.metabolites[k] += delta_m[k];
      nxt[1][idx] = p;
    }
    endJacobi(1);
  }

  // ---- metabolite reactions (extension; the reference only diffuses) --------
  // Per active metabolic particle, in table order: the fraction
  // min(rate * n, 1) of m[from] moves to m[to]; n = 1, or the number of
  // active catalysts (plain particles of catalyst_type) within
  // catalyst_distance.  Each particle changes only itself.
  void react() {
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < nActive[1]; idx++) {
      Particle& m = cur[1][idx];
      if (!active(m)) continue;
      for (const sc_reaction& r : reactions) {
        real n = 1.0f;
        if (r.catalyst_type >= 0) {
          int count = 0;
          forNeighbors(0, m.pos, [&](int j, const Particle& tp) {
            (void)j;
            if (!active(tp) || tp.type != r.catalyst_type) return;
            float3 d = tp.pos - m.pos;
            if (outside(d, r.catalyst_distance)) return;
            if (PLANT(61) || norm(d) <= r.catalyst_distance) count++;
          });
          n = (real)count;
        }
        real f = PLANT(62) ? r.rate * n : scm::fmin_(r.rate * n, 1.0f);
        real flux = m.metabolites[r.from] * f;
        m.metabolites[r.from] = m.metabolites[r.from] - flux;
        if (!PLANT(63)) m.metabolites[r.to] = m.metabolites[r.to] + flux;
      }
    }
  }

  // ---- metabolites.cuh:18-88 ------------------------------------------------
  void removeInterfering() {
    beginJacobi(1);
    for (int idx = 0; idx < nActive[1]; idx++) {
      Particle p = cur[1][idx];
      if (!active(p)) { nxt[1][idx] = p; continue; }
      forNeighbors(0, p.pos, [&](int j, const Particle& tp) {
        (void)j;
        if (!active(tp) || (tp.type != SC_TYPE_LIPID && !PLANT(12))) return;
        float3 delta = tp.pos - p.pos;
        if (outside(delta, P.max_interaction_distance)) return;
        real dist = norm(delta);
        real interferenceDist = PLANT(13) ? p.radius + tp.radius
                              : PLANT(11) ? (real)((double)(p.radius + p.radius) + 0.001)
                                          : (real)((double)(p.radius + tp.radius) + 0.001);
        if (PLANT(14) ? dist < interferenceDist : dist <= interferenceDist) p.flags = p.flags & ~SC_FLAG_ACTIVE;
      });
      nxt[1][idx] = p;
    }
    endJacobi(1);
  }

  // ---- site bonds (extension of interactions.cuh:216-346) -----------------
  // Rows with bond_probability > 0 bond by contact of the second partner's
  // centre with the first partner's site (first = participantOrder), a
  // Philox draw shared by both partners, and a mutual choice: each particle
  // picks per group the candidate with the lowest (partner id, row id), and
  // the bond is added to both only when the two picked each other.
  static constexpr int kSiteChoices = 4;
  static constexpr uint32_t kSiteKey = 0x9E3779B9u;   // added to the seed's high word: a stream of its own
  float3 minImage(float3 d) const {
    if (P.boundary != SC_BOUNDARY_PERIODIC || PLANT(53)) return d;
    const real L = P.sim_size, h = 0.5f * P.sim_size;
    auto w = [&](real x) { return x > h ? x - L : (x < -h ? x + L : x); };
    return mk3(w(d.x), w(d.y), w(d.z));
  }
  // distance of the second partner's centre from the first partner's site,
  // from the stored positions (the same operands whichever partner asks)
  real siteDistance(const Particle& first, const Particle& second, const sc_interaction_info& pii) const {
    float3 d = minImage(second.pos - first.pos) - transform_vector(rp_of(pii), first.rot);
    return norm(d);
  }
  real siteDraw(const Particle& a, const Particle& b, int rowId, uint32_t drawStep) const {
    uint32_t lo = (uint32_t)std::min(a.id, b.id), hi = (uint32_t)std::max(a.id, b.id);
    if (PLANT(55)) { lo = (uint32_t)a.id; hi = (uint32_t)b.id; }     // planted: each side its own draw
    scm::u32x4 c = {lo, hi, (uint32_t)rowId, drawStep};
    return scm::u01_(scm::philox4x32_10(c, P.seed_lo, P.seed_hi + kSiteKey).x);
  }
  bool hasSiteRows(int type) const {
    int cs, ce;
    partnerRange(type, cs, ce);
    for (int k = cs; k < ce; k++)
      if (partner[k].bond_probability > 0.0f) return true;
    return false;
  }
  void siteChoose(uint32_t drawStep) {
    const int n = nActive[0];
    siteChoice.assign((size_t)capacity(0) * 2 * kSiteChoices, -1);
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < n; idx++) {
      const Particle& p = cur[0][idx];
      if (!active(p) || !hasSiteRows(p.type)) continue;
      int cs, ce;
      partnerRange(p.type, cs, ce);
      int bgroup[kSiteChoices], brow[kSiteChoices], bslot[kSiteChoices];
      uint64_t bkey[kSiteChoices];
      int nb = 0;
      const int room = SC_MAX_BONDS - p.nActiveInteractions;
      forNeighbors(0, p.pos, [&](int j, const Particle& tps) {
        if (j == idx) return;
        const Particle& tp = cur[0][j];
        if (!active(tp) || tp.nActiveInteractions >= SC_MAX_BONDS) return;
        if (outside(tps.pos - p.pos, P.max_interaction_distance)) return;
        for (int k = cs; k < ce; k++) {
          const sc_interaction_info& pii = partner[k];
          if (!(pii.bond_probability > 0.0f) || pii.only_via_transition) continue;
          if (!((p.type == pii.first_partner_type && tp.type == pii.second_partner_type) ||
                (p.type == pii.second_partner_type && tp.type == pii.first_partner_type)))
            continue;
          if (groupActive(p, pii.group) || groupActive(tp, pii.group) || !stateMatch(pii, p, tp)) continue;
          const bool ord = participantOrder(p, tp);
          const real d = ord ? siteDistance(p, tp, pii) : siteDistance(tp, p, pii);
          if (!(d <= pii.contact_distance)) continue;
          if (!(siteDraw(p, tp, pii.id, drawStep) < pii.bond_probability)) continue;
          const uint64_t key = ((uint64_t)(uint32_t)tp.id << 32) | (uint32_t)pii.id;
          int q = 0;
          while (q < nb && bgroup[q] != pii.group) q++;
          if (q < nb) {
            if (PLANT(52) ? key > bkey[q] : key < bkey[q]) { bkey[q] = key; brow[q] = pii.id; bslot[q] = j; }
          } else if (nb < kSiteChoices && nb < room) {
            bgroup[nb] = pii.group; bkey[nb] = key; brow[nb] = pii.id; bslot[nb] = j;
            nb++;
          }
        }
      });
      int32_t* out = &siteChoice[(size_t)idx * 2 * kSiteChoices];
      for (int q = 0; q < nb; q++) { out[2 * q] = brow[q]; out[2 * q + 1] = bslot[q]; }
    }
  }
  void siteCommit() {
    const int n = nActive[0];
    SC_ORACLE_PARALLEL
    for (int idx = 0; idx < n; idx++) {
      const int32_t* mine = &siteChoice[(size_t)idx * 2 * kSiteChoices];
      for (int q = 0; q < kSiteChoices && mine[2 * q] >= 0; q++) {
        const int row = mine[2 * q], j = mine[2 * q + 1];
        const int32_t* theirs = &siteChoice[(size_t)j * 2 * kSiteChoices];
        bool mutual = false;
        for (int r = 0; r < kSiteChoices; r++)
          if (theirs[2 * r] == row && theirs[2 * r + 1] == idx) mutual = true;
        if (!mutual && !PLANT(51)) continue;
        const int group = (row >= 0 && (size_t)row < flat.size()) ? flat[row].group : 0;
        append_bond(cur[0][idx], row, group, cur[0][j].id);
      }
    }
  }
  void siteBonds(int stepArg) {
    bool any = false;
    for (const sc_interaction_info& r : partner) any = any || r.bond_probability > 0.0f;
    if (!any) return;
    const uint32_t drawStep = stepArg == -1 ? 0xffffffffu : (uint32_t)step;
    siteChoose(drawStep);
    siteCommit();
  }

  bool stateMatch(const sc_interaction_info& pii, const Particle& p, const Particle& tp) const {
    int8_t fs = (int8_t)pii.first_partner_state, ss = (int8_t)pii.second_partner_state;
    int8_t ps = (int8_t)p.state, ts = (int8_t)tp.state;
    return participantOrder(p, tp) ? ((fs < 0 || fs == ps) && (ss < 0 || ss == ts))
                                   : ((ss < 0 || ss == ps) && (fs < 0 || fs == ts));
  }
  static float4 ro_of(const sc_interaction_info& pii) {
    return mk4(pii.relative_orientation[0], pii.relative_orientation[1], pii.relative_orientation[2],
               pii.relative_orientation[3]);
  }
  static float3 rp_of(const sc_interaction_info& pii) {
    return mk3(pii.relative_position[0], pii.relative_position[1], pii.relative_position[2]);
  }
  static bool groupActive(const Particle& p, int group) {
    for (int k = 0; k < p.nActiveInteractions; k++)
      if (p.interactions[k].group == group) return true;
    return false;
  }
  void partnerRange(int type, int& s, int& e) const {
    s = e = 0;
    if (type >= 0 && 2 * (size_t)type + 1 < pidx.size()) { s = pidx[2 * type]; e = pidx[2 * type + 1]; }
  }

  // ---- interactions.cuh:216-346 ---------------------------------------------
  void complexify(int stepArg) {
    (void)stepArg;
    beginJacobi(0);
    for (int idx = 0; idx < nActive[0]; idx++) {
      Particle p = cur[0][idx];
      if (!active(p)) { nxt[0][idx] = p; continue; }
      float3 up = transform_vector(VEC_UP(), p.rot);
      int cs, ce;
      partnerRange(p.type, cs, ce);
      if (cs < ce) {
        forNeighbors(0, p.pos, [&](int j, const Particle& tp) {
          if (j == idx || !active(tp)) return;
          float3 delta = tp.pos - p.pos;
          if (outside(delta, P.max_interaction_distance)) return;
          float3 tup = transform_vector(VEC_UP(), tp.rot);
          for (int k = cs; k < ce; k++) {
            if (p.nActiveInteractions >= SC_MAX_BONDS) continue;
            const sc_interaction_info& pii = partner[k];
            if (!((p.type == pii.first_partner_type && tp.type == pii.second_partner_type) ||
                  (p.type == pii.second_partner_type && tp.type == pii.first_partner_type)) ||
                pii.only_via_transition)
              continue;
            if (pii.bond_probability > 0.0f) continue;      // site-contact rows bond in siteBonds
            if (pii.group >= 0 && (groupActive(p, pii.group) || groupActive(tp, pii.group))) continue;
            bool ord = participantOrder(p, tp);
            float4 ro = ro_of(pii);
            float3 rp = rp_of(pii);
            real oa = ord ? dot(transform_vector(up, ro), tup) : dot(transform_vector(tup, ro), up);
            real rpa = ord ? dot(transform_vector(rp, p.rot), delta) / normsq(delta)
                            : dot(transform_vector(rp, tp.rot), -delta) / normsq(delta);
            if (oa > pii.orientation_alignment_threshold &&
                scm::dabs_((double)rpa - 1.0) < (1.0 - (double)pii.position_alignment_threshold) &&
                stateMatch(pii, p, tp))
              append_bond(p, pii.id, pii.group, tp.id);
          }
        });
      }
      nxt[0][idx] = p;
    }
    endJacobi(0);
  }

  // ---- interactions.cuh:349-555 ---------------------------------------------
  void transitionInteractions(int stepArg) {
    beginJacobi(0);
    for (int idx = 0; idx < nActive[0]; idx++) {
      Particle p = cur[0][idx];
      if (!active(p)) { nxt[0][idx] = p; continue; }
      float3 up = transform_vector(VEC_UP(), p.rot);
      int sourceTransitionIndex = -1;
      sc_interaction_info targetTransition;
      std::memset(&targetTransition, 0, sizeof(targetTransition));
      if (p.nActiveInteractions > 0) {
        forNeighbors(0, p.pos, [&](int j, const Particle& tp) {
          if (j == idx || !active(tp)) return;
          float3 delta = tp.pos - p.pos;
          if (outside(delta, P.max_interaction_distance)) return;
          float3 tup = transform_vector(VEC_UP(), tp.rot);
          for (int k = 0; k < p.nActiveInteractions; k++) {
            Bond b = p.interactions[k];
            if (b.type < 10) continue;
            if (b.partnerId != tp.id) continue;
            sc_interaction_info pii = info(b.type);
            if (!(pii.transition_to >= 0 || pii.break_after_transition)) continue;
            bool ord = participantOrder(p, tp);
            float4 ro = ro_of(pii);
            float3 rp = rp_of(pii);
            real oa = ord ? dot(transform_vector(up, ro), tup) : dot(transform_vector(tup, ro), up);
            real rpa = ord ? dot(transform_vector(rp, p.rot), delta) / normsq(delta)
                            : dot(transform_vector(rp, tp.rot), -delta) / normsq(delta);
            bool aligned = pii.break_after_transition || !pii.wait_for_alignment ||
                           (oa > pii.orientation_alignment_threshold &&
                            scm::dabs_((double)rpa - 1.0) < (1.0 - (double)pii.position_alignment_threshold));
            bool stateOk = !pii.wait_for_state || stateMatch(pii, p, tp);
            if (!(aligned && stateOk)) continue;
            if (pii.transition_to >= 0) {
              sc_interaction_info tpii = info(pii.transition_to);
              if (tpii.group >= 0 && tpii.group != pii.group &&
                  (groupActive(p, tpii.group) || groupActive(tp, tpii.group)))
                continue;
              int8_t tfs = (int8_t)tpii.first_partner_state, tss = (int8_t)tpii.second_partner_state;
              if (!tpii.set_state || !((p.type == tpii.first_partner_type && tfs >= 0) ||
                                       (p.type == tpii.second_partner_type && tss >= 0))) {
                p.interactions[k].type = tpii.id;
                p.interactions[k].group = tpii.group;
              } else {
                sourceTransitionIndex = k;
                targetTransition = tpii;
              }
            } else if (pii.break_after_transition) {
              for (int l = k; l < p.nActiveInteractions - 1; l++) p.interactions[l] = p.interactions[l + 1];
              p.nActiveInteractions--;
              if (sourceTransitionIndex > k) sourceTransitionIndex--;
              k--;
            }
          }
        });
      }
      if (sourceTransitionIndex >= 0) {
        p.interactions[sourceTransitionIndex].type = targetTransition.id;
        p.interactions[sourceTransitionIndex].group = targetTransition.group;
        if (targetTransition.set_state) {
          int8_t tfs = (int8_t)targetTransition.first_partner_state, tss = (int8_t)targetTransition.second_partner_state;
          if (p.type == targetTransition.first_partner_type && tfs >= 0) { p.state = tfs; p.lastStateChangeStep = stepArg; }
          if (p.type == targetTransition.second_partner_type && tss >= 0) { p.state = tss; p.lastStateChangeStep = stepArg; }
        }
        if (targetTransition.polymerize >= 0) {
          sc_interaction_info poly = info(targetTransition.polymerize);
          if (poly.first_partner_type == p.type || poly.second_partner_type == p.type)
            p.scheduledPolymerization = targetTransition.polymerize;
        }
      }
      nxt[0][idx] = p;
    }
    endJacobi(0);
  }

  // ---- interactions.cuh:558-740, serial in slot order -----------------------
  void polymerize(int stepArg) {
    int stepStart = nActive[0];
    int lastActive = nActive[0] - 1;
    std::vector<Particle>& A = cur[0];
    for (int idx = 0; idx < stepStart; idx++) {
      Particle p = A[idx];
      if (!active(p) || p.scheduledPolymerization < 0) continue;
      sc_interaction_info poly = info(p.scheduledPolymerization);
      int activeIdx = -1;
      for (int k = 0; k < p.nActiveInteractions; k++)
        if (p.interactions[k].group == poly.group) { activeIdx = k; break; }
      int prevIdx = -1;
      Particle prev;
      std::memset(&prev, 0, sizeof(prev));
      if (activeIdx >= 0) {
        int target = p.interactions[activeIdx].partnerId;
        forStencil(p.pos, [&](uint32_t cell, real, real, real) {
          uint32_t s = ranges[0][2 * (size_t)cell], e = ranges[0][2 * (size_t)cell + 1];
          for (uint32_t j = s; j < e; j++) {
            if ((int)j == idx) continue;
            const Particle& tp = A[j];
            if (!active(tp)) continue;
            if (tp.id == target) { prevIdx = (int)j; prev = tp; }
          }
        });
      }
      int newType = (poly.first_partner_type == p.type) ? poly.second_partner_type : poly.first_partner_type;
      int newState = (int8_t)((poly.first_partner_type == p.type) ? poly.second_partner_state : poly.first_partner_state);
      if (newType >= 0) {
        int newIdx = ++lastActive;
        if (newIdx < P.num_particles) {
          int newId = nextParticleId++;
          sc_type_info ti = {2.5f, 0};
          if (newType < (int)types.size()) ti = types[newType];
          else { ti.radius = 0.0f; ti.hydrophobic = 0; }
          Particle np;
          std::memset(&np, 0, sizeof(np));
          np.id = newId; np.type = newType; np.flags = SC_FLAG_ACTIVE; np.state = 0;
          np.lastStateChangeStep = -1; np.radius = ti.radius; np.hydrophobic = (int8_t)ti.hydrophobic;
          float3 rp = rp_of(poly);
          float4 ro = ro_of(poly);
          np.pos = p.pos + transform_vector(p.type < newType ? rp : -rp, p.rot);
          np.rot = qmul(p.rot, p.type < newType ? ro : inverse(ro));
          np.velocity = VEC_ZERO(); np.posNoise = VEC_ZERO();
          np.angularVelocity = QUAT_ID(); np.angularNoise = QUAT_ID();
          np.debugVector = mk4(0, 0, 0, 0);
          np.scheduledPolymerization = -1;
          if (newState >= 0) { np.state = newState; np.lastStateChangeStep = stepArg; }
          append_bond(np, p.scheduledPolymerization, poly.group, p.id);
          if (prevIdx >= 0) {
            for (int k = 0; k < prev.nActiveInteractions; k++) {
              if (prev.interactions[k].group == poly.group) {
                for (int l = k; l < prev.nActiveInteractions - 1; l++) prev.interactions[l] = prev.interactions[l + 1];
                prev.nActiveInteractions--;
                break;
              }
            }
            append_bond(prev, SC_INTERACTION_TYPE_CHAIN, SC_GROUP_FORWARD, np.id);
            A[prevIdx] = prev;
            append_bond(np, SC_INTERACTION_TYPE_CHAIN, SC_GROUP_BACKWARD, prev.id);
            p.interactions[activeIdx].type = p.scheduledPolymerization;
            p.interactions[activeIdx].partnerId = np.id;
          } else {
            append_bond(p, p.scheduledPolymerization, poly.group, np.id);
          }
          A[newIdx] = np;
          nActive[0]++;
        }
      } else {
        if (prevIdx >= 0) {
          for (int k = 0; k < prev.nActiveInteractions; k++) {
            if (prev.interactions[k].group == poly.group) {
              for (int l = k; l < prev.nActiveInteractions - 1; l++) prev.interactions[l] = prev.interactions[l + 1];
              prev.nActiveInteractions--;
              break;
            }
          }
          A[prevIdx] = prev;
        }
        if (activeIdx >= 0) {
          for (int l = activeIdx; l < p.nActiveInteractions - 1; l++) p.interactions[l] = p.interactions[l + 1];
          p.nActiveInteractions--;
        }
      }
      p.scheduledPolymerization = -1;
      A[idx] = p;
    }
  }

  // ---- interactions.cuh:743-822 (not launched by the reference's loop,
  // main.cu:618-631; in the step when the config enables it) ----------------
  // A particle takes the state of a bonded partner (bond type >= 10 whose
  // interaction propagates state) that changed state at a later step;
  // partners in stencil order, each compared with the state taken so far.
  void propagateState() {
    beginJacobi(0);
    for (int idx = 0; idx < nActive[0]; idx++) {
      Particle p = cur[0][idx];
      if (!active(p)) { nxt[0][idx] = p; continue; }
      forNeighbors(0, p.pos, [&](int j, const Particle& tp) {
        if (j == idx || !active(tp)) return;
        float3 delta = tp.pos - p.pos;
        if (outside(delta, P.max_interaction_distance)) return;
        for (int k = 0; k < p.nActiveInteractions; k++) {
          Bond b = p.interactions[k];
          if (b.type < 10 || b.partnerId != tp.id) continue;
          if (!info(b.type).propagate_state) continue;
          if (tp.lastStateChangeStep > p.lastStateChangeStep) {
            p.lastStateChangeStep = tp.lastStateChangeStep;
            p.state = tp.state;
          }
        }
      });
      nxt[0][idx] = p;
    }
    endJacobi(0);
  }

  void runStage(int stage, int arg) {
    switch (stage) {
      case SC_STAGE_SORT: updateGridAndSort(0); updateGridAndSort(1); break;
      case SC_STAGE_MOVE:
        moveKernel(0); brownian(0, P.movement_noise_scale);
        moveKernel(1); brownian(1, P.metabolite_movement_noise_scale);
        break;
      case SC_STAGE_COORDINATE: coordinateNoise((real)arg / P.noise_coordination_steps); break;
      case SC_STAGE_APPLY_NOISE: applyNoise(0); applyNoise(1); break;
      case SC_STAGE_RELAX: relax(); break;
      case SC_STAGE_RELAX_METABOLIC: relaxMetabolicParticles(); break;
      case SC_STAGE_RELAX_PARTNERS: relaxMetabolicParticlePartners(); break;
      case SC_STAGE_APPLY_VELOCITIES: applyVelocities(0, 0.45f); applyVelocities(1, 0.45f); break;
      case SC_STAGE_COMPLEXIFY: complexify(arg); siteBonds(arg); break;
      case SC_STAGE_TRANSITION: transitionInteractions(arg); break;
      case SC_STAGE_POLYMERIZE: polymerize(arg); break;
      case SC_STAGE_DIFFUSE: diffuseMetabolites(); break;
      case SC_STAGE_REMOVE_INTERFERING: removeInterfering(); break;
      case SC_STAGE_END_STEP: step++; break;
      case SC_STAGE_PROPAGATE_STATE: propagateState(); break;
      case SC_STAGE_REACT: react(); break;
      default: break;
    }
  }

  // main.cu:347-649
  void fullStep() {
    runStage(SC_STAGE_SORT, 0);
    runStage(SC_STAGE_MOVE, 0);
    for (int j = 0; j < P.noise_coordination_steps; j++) runStage(SC_STAGE_COORDINATE, j);
    runStage(SC_STAGE_APPLY_NOISE, 0);
    for (int j = 0; j < P.relaxation_steps; j++) {
      runStage(SC_STAGE_RELAX, j);
      if (j % 2 == 0) { runStage(SC_STAGE_RELAX_METABOLIC, j); runStage(SC_STAGE_RELAX_PARTNERS, j); }
    }
    runStage(SC_STAGE_APPLY_VELOCITIES, 0);
    runStage(SC_STAGE_COMPLEXIFY, step);
    runStage(SC_STAGE_TRANSITION, step);
    runStage(SC_STAGE_POLYMERIZE, step);
    if (P.propagate_state) runStage(SC_STAGE_PROPAGATE_STATE, 0);
    for (int j = 0; j < P.metabolite_diffusion_steps; j++) runStage(SC_STAGE_DIFFUSE, j);
    if (!reactions.empty()) runStage(SC_STAGE_REACT, 0);
    runStage(SC_STAGE_END_STEP, 0);
  }
};

extern "C" {

orc_sim* orc_create(const sc_params* params, const sc_tables* tables, const sc_counts* counts,
                    const sc_particle* plain, size_t n_plain, const sc_particle* metabolic, size_t n_metabolic) {
  orc_sim* s = new orc_sim();
  s->P = *params;
  // config.cuh:66-74
  int n = (int)std::ceil(s->P.sim_size / s->P.grid_cell_size);
  int bits = 1, p2 = 2;
  while (p2 < n) { p2 *= 2; bits++; }
  s->nGridCells = p2;
  s->nGridCellsBits = bits;
  s->nPer = (int)(s->P.sim_size / s->P.grid_cell_size);
  if (tables) {
    s->types.assign(tables->types, tables->types + tables->n_types);
    s->flat.assign(tables->flat, tables->flat + tables->n_flat);
    s->partner.assign(tables->partner, tables->partner + tables->n_partner);
    s->pidx.assign(tables->partner_index, tables->partner_index + tables->n_partner_index);
    if (tables->n_reactions > 0 && tables->reactions)
      s->reactions.assign(tables->reactions, tables->reactions + tables->n_reactions);
  }
  size_t cells = (size_t)p2 * p2 * p2;
  for (int pop = 0; pop < 2; pop++) {
    int cap = s->capacity(pop);
    Particle z;
    std::memset(&z, 0, sizeof(z));
    s->cur[pop].assign(cap, z);
    s->nxt[pop].assign(cap, z);
    s->ranges[pop].assign(2 * cells, 0u);
    const sc_particle* src = pop == 0 ? plain : metabolic;
    size_t ns = pop == 0 ? n_plain : n_metabolic;
    for (size_t i = 0; i < ns && i < (size_t)cap; i++) s->cur[pop][i] = from_sc(src[i]);
    s->nSorted[pop] = 0;
  }
  s->nActive[0] = counts->n_active;
  s->nActive[1] = counts->n_active_metabolic;
  s->nextParticleId = counts->next_particle_id;
  s->step = counts->step;
  return s;
}

void orc_destroy(orc_sim* s) { delete s; }

void orc_step(orc_sim* s, int n) {
  for (int i = 0; i < n; i++) s->fullStep();
}

int orc_run_stage(orc_sim* s, int stage, int arg) {
  if (stage < 0 || stage >= SC_NUM_STAGES) return -1;
  s->runStage(stage, arg);
  return 0;
}

void orc_snapshot(const orc_sim* s, sc_counts* counts, sc_particle* plain, size_t plain_cap,
                  sc_particle* metabolic, size_t metabolic_cap) {
  if (counts) {
    counts->n_active = s->nActive[0];
    counts->n_active_metabolic = s->nActive[1];
    counts->next_particle_id = s->nextParticleId;
    counts->step = s->step;
  }
  for (size_t i = 0; plain && i < plain_cap && i < s->cur[0].size(); i++) to_sc(s->cur[0][i], plain[i], false);
  for (size_t i = 0; metabolic && i < metabolic_cap && i < s->cur[1].size(); i++)
    to_sc(s->cur[1][i], metabolic[i], true);
}

size_t orc_debug_export(const orc_sim* s, int what, int population, void* dst, size_t cap) {
  if (population < 0 || population > 1) return 0;
  int n = s->nSorted[population];
  if (what == SC_EXPORT_KEYS) {
    size_t need = (size_t)n * 4;
    if (dst && cap >= need) std::memcpy(dst, s->sortedKeys[population].data(), need);
    return need;
  }
  if (what == SC_EXPORT_PERMUTATION) {
    size_t need = (size_t)n * 4;
    if (dst && cap >= need) std::memcpy(dst, s->perm[population].data(), need);
    return need;
  }
  if (what == SC_EXPORT_RANGES) {
    std::vector<uint32_t> out;
    const std::vector<uint32_t>& R = s->ranges[population];
    for (size_t c = 0; c < R.size() / 2; c++)
      if (R[2 * c + 1] > R[2 * c]) { out.push_back((uint32_t)c); out.push_back(R[2 * c]); out.push_back(R[2 * c + 1]); }
    size_t need = out.size() * 4;
    if (dst && cap >= need) std::memcpy(dst, out.data(), need);
    return need;
  }
  if (what == SC_EXPORT_NEIGHBORS) {
    std::vector<int32_t> out;
    for (int i = 0; i < s->nActive[population]; i++) {
      const Particle& p = s->cur[population][i];
      if (!active(p)) continue;
      s->forNeighbors(population, p.pos, [&](int j, const Particle& tp) {
        if (j == i || !active(tp)) return;
        if (s->outside(tp.pos - p.pos, s->P.max_interaction_distance)) return;
        out.push_back(i); out.push_back(j);
      });
    }
    size_t need = out.size() * 4;
    if (dst && cap >= need) std::memcpy(dst, out.data(), need);
    return need;
  }
  return 0;
}

// dynamics.cuh:824-844, types.cuh:238-294, storage.cuh:29-39
size_t orc_frame_size(const orc_sim* s) {
  return 4 + (size_t)s->P.num_particles * SC_REDUCED_PARTICLE_BYTES + 4 +
         (size_t)s->P.num_metabolic_particles * SC_REDUCED_METABOLIC_BYTES;
}

size_t orc_emit_frame(const orc_sim* s, void* dst, size_t cap) {
  size_t need = orc_frame_size(s);
  if (!dst || cap < need) return need;
  unsigned char* o = (unsigned char*)dst;
  std::memset(o, 0, need);
  for (int pop = 0; pop < 2; pop++) {
    uint32_t capn = (uint32_t)s->capacity(pop);
    std::memcpy(o, &capn, 4);
    o += 4;
    size_t rec = pop == 0 ? SC_REDUCED_PARTICLE_BYTES : SC_REDUCED_METABOLIC_BYTES;
    for (int i = 0; i < s->nActive[pop]; i++) {
      const Particle& p = s->cur[pop][i];
      unsigned char* r = o + (size_t)i * rec;
      int32_t head[3] = {p.id, p.type, p.flags | (((int)p.state & 0x0000000F) << 16)};
      std::memcpy(r, head, 12);
      real f[12] = {p.radius, p.pos.x, p.pos.y, p.pos.z, p.rot.x, p.rot.y, p.rot.z, p.rot.w,
                     p.debugVector.x, p.debugVector.y, p.debugVector.z, p.debugVector.w};
      for (int k = 0; k < 12; k++) { uint16_t h = scm::f2h_(f[k]); std::memcpy(r + 12 + 2 * k, &h, 2); }
      if (pop == 1)
        for (int k = 0; k < SC_REDUCED_NUM_METABOLITES; k++) {
          uint16_t h = scm::f2h_(p.metabolites[k]);
          std::memcpy(r + 36 + 2 * k, &h, 2);
        }
    }
    o += (size_t)capn * rec;
  }
  return need;
}

// ---- test hooks: primitives and vector math of this (g++) build -------------
float orc_eval_primitive1(int fn, float x) {
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
      float4 q = random_rotation(scm::u01_(b), scm::u01_(b * 747796405u), scm::u01_(b * 2891336453u));
      float3 u = transform_vector(VEC_UP(), q);
      return u.x + u.y * 3.0f + u.z * 7.0f;
    }
    default: return 0.0f;
  }
}

void orc_eval_primitive(int fn, const float* x, float* y, size_t n) {
  for (size_t i = 0; i < n; i++) y[i] = orc_eval_primitive1(fn, x[i]);
}

void orc_philox(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4, size_t n) {
  for (size_t i = 0; i < n; i++) {
    scm::u32x4 c = {ctr4[4 * i], ctr4[4 * i + 1], ctr4[4 * i + 2], ctr4[4 * i + 3]};
    scm::u32x4 r = scm::philox4x32_10(c, key2[2 * i], key2[2 * i + 1]);
    out4[4 * i] = r.x; out4[4 * i + 1] = r.y; out4[4 * i + 2] = r.z; out4[4 * i + 3] = r.w;
  }
}

// Vector/quaternion helpers of math.cuh: in = a(4) b(4) t, out(4).
// 0 transform_vector(a.xyz, b)   1 qmul(a, b)   2 inverse(a)   3 slerp(a, b, t)
// 4 quaternionFromTo(a.xyz, b.xyz)   5 quaternion(a.xyz, t)   6 angle(a.xyz, b.xyz)
// 7 safeCross(a.xyz, b.xyz)   8 normalize(a.xyz)
void orc_eval_math(int fn, const float* in, float* out, size_t n) {
  for (size_t i = 0; i < n; i++) {
    const float* v = in + 9 * i;
    float* o = out + 4 * i;
    float4 a = mk4(v[0], v[1], v[2], v[3]), b = mk4(v[4], v[5], v[6], v[7]);
    float t = v[8];
    float3 a3 = mk3(a.x, a.y, a.z), b3 = mk3(b.x, b.y, b.z);
    float4 r = mk4(0, 0, 0, 0);
    switch (fn) {
      case 0: { float3 x = transform_vector(a3, b); r = mk4(x.x, x.y, x.z, 0); break; }
      case 1: r = qmul(a, b); break;
      case 2: r = inverse(a); break;
      case 3: r = slerp(a, b, t); break;
      case 4: r = quaternionFromTo(a3, b3); break;
      case 5: r = quaternion(a3, t); break;
      case 6: r = mk4(angle(a3, b3), 0, 0, 0); break;
      case 7: { float3 x = safeCross(a3, b3); r = mk4(x.x, x.y, x.z, 0); break; }
      case 8: { float3 x = normalize(a3); r = mk4(x.x, x.y, x.z, 0); break; }
      default: break;
    }
    o[0] = r.x; o[1] = r.y; o[2] = r.z; o[3] = r.w;
  }
}

}  // extern "C"
