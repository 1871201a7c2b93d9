// Host-side setup: config.json parsing (spatial_cell_sim/src/config.cuh:46-75),
// the table loaders (chemistry.cuh:10-51, interactions.cuh:17-145) and
// synthetic population generators that replace the hard-coded scene of
// setup/setup.cuh.  Positions and orientations come from Philox4x32-10 keyed
// by the config seed, so a setup document describes its scene exactly.
#include "sc_setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "sc_fmath.h"
#include "sc_json.h"
#include "sc_vec.h"

namespace scs {

namespace {

const uint32_t kSetupStream = 0x53455450u;  // Philox counter word for scene generation

struct Rng {
  uint32_t k0, k1, stream, pop;
  scm::u32x4 block(uint32_t index, uint32_t blk) const {
    scm::u32x4 c = {index, pop, stream, blk};
    return scm::philox4x32_10(c, k0, k1);
  }
};

float u01(uint32_t x) { return scm::u01_(x); }

void shoemake(float u, float v, float w, float q[4]) {
  const float PI = 3.1415926535f;
  float su = scm::sqrt_(u), su1 = scm::sqrt_(1 - u);
  q[0] = su1 * scm::sin_(2 * PI * v);
  q[1] = su1 * scm::cos_(2 * PI * v);
  q[2] = su * scm::sin_(2 * PI * w);
  q[3] = su * scm::cos_(2 * PI * w);
}

void unit_vector(float u, float v, float d[3]) {
  const float PI = 3.1415926535f;
  float z = 2.0f * u - 1.0f;
  float r = scm::sqrt_(std::max(0.0f, 1.0f - z * z));
  d[0] = r * scm::cos_(2 * PI * v);
  d[1] = r * scm::sin_(2 * PI * v);
  d[2] = z;
}

float clampL(float x, float L) { return std::min(std::max(x, 0.0f), L); }

// fillParticlesSphere's points (setup/particle_setup.cuh:174-191): a
// Fibonacci spiral, in double, cast to float
void fibonacci_sphere(int n, std::vector<scv::v3>& out) {
  const double gr = (std::sqrt(5.0) + 1.0) / 2.0;
  const double ga = (2.0 - gr) * (2.0 * (double)3.1415926535f);
  out.clear();
  for (int i = 1; i <= n; i++) {
    const double lat = std::asin(-1.0 + 2.0 * (double)i / (n + 1));
    const double lon = ga * i;
    out.push_back(scv::m3((float)(std::cos(lon) * std::cos(lat)), (float)(std::sin(lon) * std::cos(lat)),
                          (float)std::sin(lat)));
  }
}

// three numbers of a JSON array (missing entries: d)
bool read3(const scj::Value* v, double d, double out[3]) {
  for (int k = 0; k < 3; k++) out[k] = d;
  if (!v) return false;
  for (int k = 0; k < 3 && k < (int)v->size(); k++) out[k] = (*v)[k].as_double();
  return true;
}

sc_interaction_info parse_interaction(const scj::Value& c) {
  // interactions.cuh:31-50 defaults; a key present with null reads as 0/false
  sc_interaction_info r;
  std::memset(&r, 0, sizeof(r));
  r.id = c.get_int("id", 0);
  r.group = c.get_int("group", -1);
  r.first_partner_type = c.get_int("firstPartnerType", 0);
  r.first_partner_state = (int8_t)c.get_int("firstPartnerState", -1);
  r.second_partner_type = c.get_int("secondPartnerType", 0);
  r.second_partner_state = (int8_t)c.get_int("secondPartnerState", -1);
  r.propagate_state = c.get_bool("propagateState", false);
  r.set_state = c.get_bool("setState", true);
  r.wait_for_state = c.get_bool("waitForState", true);
  r.wait_for_alignment = c.get_bool("waitForAlignment", true);
  r.only_via_transition = c.get_bool("onlyViaTransition", false);
  r.transition_to = c.get_int("transitionTo", -1);
  r.polymerize = c.get_int("polymerize", -1);
  r.break_after_transition = c.get_bool("breakAfterTransition", false);
  const scj::Value* ro = c.find("relativeOrientation");
  const scj::Value* rp = c.find("relativePosition");
  for (int k = 0; k < 4; k++) r.relative_orientation[k] = (ro && ro->size() > (size_t)k) ? (*ro)[k].as_float() : 0.0f;
  for (int k = 0; k < 3; k++) r.relative_position[k] = (rp && rp->size() > (size_t)k) ? (*rp)[k].as_float() : 0.0f;
  r.orientation_alignment_threshold = c.get_float("orientationAlignmentThreshold", 0.8f);
  r.position_alignment_threshold = c.get_float("positionAlignmentThreshold", 0.8f);
  r.bond_probability = c.get_float("bondProbability", 0.0f);     // extension: site-contact bonds
  r.contact_distance = c.get_float("contactDistance", 1.0f);
  return r;
}

bool load_doc(const scj::Value& setup, const char* inline_key, const char* file_key, scj::Value& out,
              std::string& err) {
  if (const scj::Value* v = setup.find(inline_key)) { out = *v; return true; }
  std::string path = setup.get_str(file_key, "");
  if (path.empty()) { out = scj::Value(); out.kind = scj::Value::Array; return true; }
  std::string text;
  if (!scj::read_file(path, text)) { err = std::string("cannot read ") + file_key + " '" + path + "'"; return false; }
  if (!scj::parse(text, out, err)) { err = path + ": " + err; return false; }
  return true;
}

void link_serially(std::vector<sc_particle>& a, int start, int end, int type) {
  // setup/interaction_setup.cuh:13-43
  for (int i = start + 1; i < end; i++) {
    sc_particle& p = a[i];
    sc_particle& q = a[i - 1];
    if (q.n_bonds < SC_MAX_BONDS) {
      q.bonds[q.n_bonds][0] = type; q.bonds[q.n_bonds][1] = SC_GROUP_FORWARD; q.bonds[q.n_bonds][2] = p.id;
      q.n_bonds++;
    }
    if (p.n_bonds < SC_MAX_BONDS) {
      p.bonds[p.n_bonds][0] = type; p.bonds[p.n_bonds][1] = SC_GROUP_BACKWARD; p.bonds[p.n_bonds][2] = q.id;
      p.n_bonds++;
    }
  }
}

}  // namespace

sc_particle default_particle(int id, int type, float radius, int hydrophobic) {
  sc_particle p;
  std::memset(&p, 0, sizeof(p));
  p.id = id;
  p.type = type;
  p.flags = SC_FLAG_ACTIVE;
  p.state = 0;
  p.last_state_change_step = -1;
  p.hydrophobic = (int8_t)hydrophobic;
  p.n_bonds = 0;
  p.scheduled_polymerization = -1;
  p.radius = radius;
  p.rot[3] = 1.0f;
  p.ang_vel[3] = 1.0f;
  p.ang_noise[3] = 1.0f;
  return p;
}

bool derive_grid(const sc_params& P, int& n_cells, int& bits, int& n_periodic, std::string& err) {
  if (!(P.sim_size > 0.0f) || !(P.grid_cell_size > 0.0f)) { err = "simSize and gridCellSize must be positive"; return false; }
  if (P.num_particles < 0 || P.num_metabolic_particles < 0) { err = "particle capacities must be >= 0"; return false; }
  if ((int64_t)P.num_particles >= (1 << 30) || (int64_t)P.num_metabolic_particles >= (1 << 30)) {
    err = "particle capacities must be below 2^30";
    return false;
  }
  int n = (int)std::ceil(P.sim_size / P.grid_cell_size);
  int b = 1, p2 = 2;
  while (p2 < n) { p2 *= 2; b++; }
  if (b > 10) { err = "grid too fine: more than 1024 cells per axis"; return false; }
  n_cells = p2;
  bits = b;
  n_periodic = (int)(P.sim_size / P.grid_cell_size);
  if (P.boundary == SC_BOUNDARY_PERIODIC) {
    if ((float)n_periodic * P.grid_cell_size != P.sim_size || n_periodic < 3) {
      err = "periodic boundary needs simSize an integer multiple (>= 3) of gridCellSize";
      return false;
    }
    if (P.max_interaction_distance > P.grid_cell_size || P.max_diffusion_distance > P.grid_cell_size) {
      err = "periodic boundary needs interaction and diffusion distances <= gridCellSize";
      return false;
    }
  } else if (P.boundary != SC_BOUNDARY_CLAMP) {
    err = "unknown boundary";
    return false;
  }
  if (P.noise_coordination_steps < 0 || P.relaxation_steps < 0 || P.metabolite_diffusion_steps < 0) {
    err = "sub-iteration counts must be >= 0";
    return false;
  }
  return true;
}

// Radius and hydrophobicity of a type; a type the table lacks reads as
// radius 0 (the reference's map lookup default-constructs its entry).
static void type_props(const HostSetup& H, int type, float& radius, int& hyd) {
  radius = 0.0f;
  hyd = 0;
  if (type >= 0 && type < (int)H.types.size()) { radius = H.types[type].radius; hyd = H.types[type].hydrophobic; }
}

// The reference scene's placements (setup/particle_setup.cuh,
// setup/interaction_setup.cuh, used by setup/setup.cuh:18-189):
//  "sphere"        fillParticlesSphere: count particles of one type on a
//                  Fibonacci sphere of radius sqrt(count / 1000 (r / 2.5)^2) 40
//                  (or "sphere_radius"), each turned from +y to its outward
//                  direction, then by "orientation" / "orientation_axis_angle";
//  "line"          fillParticlesStraightLine: the "types" list from the anchor
//                  in steps of "step", identity orientation;
//  "wrapped_chain" fillParticlesWrappedChain: the "types" list, each member
//                  turned 30 % of the way to a new random orientation and
//                  placed r(type - 1) + r(type) below the previous one along
//                  its up vector; "repeat" chains, at random positions unless
//                  anchored;
//  "complex"       instantiateComplex: complex "complex" of the complexes
//                  table about the anchor.
// The anchor is "center" (nm) or "at" (fractions of simSize, default the
// middle) plus "offset" (nm).  "link": true links each chain or line
// serially with bond type "bond_type" (linkParticlesSerially).  Positions
// are clamped to the box.  Random numbers come from the population's own
// Philox stream, so positions differ from the reference's mt19937 run.
static bool place_shape(const scj::Value& d, const std::string& shape, size_t pi, std::vector<sc_particle>& A, int cap,
                        HostSetup& out, const scj::Value& complexes, std::string& err) {
  const sc_params& P = out.P;
  const float L = P.sim_size;
  Rng rng = {P.seed_lo, P.seed_hi, kSetupStream, (uint32_t)pi};
  double at[3], off[3], c[3];
  read3(d.find("at"), 0.5, at);
  read3(d.find("offset"), 0.0, off);
  for (int q = 0; q < 3; q++) c[q] = at[q] * (double)L + off[q];
  if (d.has("center")) read3(d.find("center"), 0.0, c);
  const bool anchored = d.has("center") || d.has("at") || d.has("offset");
  const scv::v3 anchor = scv::m3((float)c[0], (float)c[1], (float)c[2]);
  scv::v4 base = scv::qid();
  if (const scj::Value* o = d.find("orientation"))
    base = scv::m4(o->size() > 0 ? (*o)[0].as_float() : 0.0f, o->size() > 1 ? (*o)[1].as_float() : 0.0f,
                   o->size() > 2 ? (*o)[2].as_float() : 0.0f, o->size() > 3 ? (*o)[3].as_float() : 1.0f);
  if (const scj::Value* o = d.find("orientation_axis_angle"))
    if (o->size() >= 4)
      base = scv::quat_axis_angle(scv::m3((*o)[0].as_float(), (*o)[1].as_float(), (*o)[2].as_float()), (*o)[3].as_float());
  std::vector<int> types;
  if (const scj::Value* t = d.find("types"))
    for (size_t k = 0; k < t->size(); k++) types.push_back((*t)[k].as_int());
  const bool link = d.get_bool("link", false);
  const int bond_type = d.get_int("bond_type", SC_INTERACTION_TYPE_CHAIN);
  auto push = [&](int type, scv::v3 pos, scv::v4 rot) -> bool {
    if ((int)A.size() >= cap) { err = "populations exceed the capacity of the array"; return false; }
    float radius;
    int hyd;
    type_props(out, type, radius, hyd);
    if (d.has("radius")) radius = d.get_float("radius", radius);
    sc_particle p = default_particle((int)A.size(), type, radius, hyd);
    p.pos[0] = clampL(pos.x, L); p.pos[1] = clampL(pos.y, L); p.pos[2] = clampL(pos.z, L);
    p.rot[0] = rot.x; p.rot[1] = rot.y; p.rot[2] = rot.z; p.rot[3] = rot.w;
    A.push_back(p);
    return true;
  };
  if (shape == "sphere") {
    const int type = d.get_int("type", 0);
    int count = d.get_int("count", 0);
    if (d.has("count_fraction")) count = (int)((double)cap * d.get_double("count_fraction", 0.0));
    float radius;
    int hyd;
    type_props(out, type, radius, hyd);
    std::vector<scv::v3> pts;
    fibonacci_sphere(count, pts);
    const float r = d.has("sphere_radius") ? d.get_float("sphere_radius", 0.0f)
                                           : (float)(std::sqrt(count / 1000.0 * std::pow(radius / 2.5, 2.0)) * 40.0);
    for (int k = 0; k < count; k++) {
      const scv::v3 pos = scv::add(anchor, scv::mul(pts[k], r));
      const scv::v3 dir = scv::nrm(scv::sub(pos, anchor));
      if (!push(type, pos, scv::qmul(scv::quat_from_to(scv::up3(), dir), base))) return false;
    }
    return true;
  }
  if (shape == "complex") {
    const int id = d.get_int("complex", 0);
    const scj::Value* cx = nullptr;
    for (size_t k = 0; k < complexes.size(); k++)
      if (complexes[k].get_int("id", -1) == id) cx = &complexes[k];
    if (!cx) { err = "populations: complex " + std::to_string(id) + " is not in the complexes table"; return false; }
    const scj::Value* parts = cx->find("participants");
    for (size_t k = 0; parts && k < parts->size(); k++) {
      const scj::Value& cp = (*parts)[k];
      double dp[3];
      read3(cp.find("position"), 0.0, dp);
      scv::v4 rot = scv::qid();
      if (const scj::Value* o = cp.find("rotation"))
        if (o->size() >= 4) rot = scv::m4((*o)[0].as_float(), (*o)[1].as_float(), (*o)[2].as_float(), (*o)[3].as_float());
      if (!push(cp.get_int("type", 0), scv::add(anchor, scv::m3((float)dp[0], (float)dp[1], (float)dp[2])), rot)) return false;
    }
    return true;
  }
  if (types.empty()) { err = "populations: shape '" + shape + "' needs a \"types\" list"; return false; }
  if (shape == "line") {
    double st[3];
    read3(d.find("step"), 0.0, st);
    const scv::v3 step = scv::m3((float)st[0], (float)st[1], (float)st[2]);
    const int start = (int)A.size();
    for (size_t k = 0; k < types.size(); k++)
      if (!push(types[k], scv::add(anchor, scv::mul(step, (float)k)), scv::qid())) return false;
    if (link) link_serially(A, start, (int)A.size(), bond_type);
    return true;
  }
  // wrapped_chain
  const int repeat = std::max(1, d.get_int("repeat", 1));
  uint32_t draw = 0;   // Philox index of this population's next draw
  auto random_rotation = [&]() {
    const scm::u32x4 b = rng.block(draw++, 1);
    return scv::shoemake(u01(b.x), u01(b.y), u01(b.z));
  };
  for (int ch = 0; ch < repeat; ch++) {
    scv::v3 pos = anchor;
    if (!anchored || repeat > 1) {
      const scm::u32x4 b = rng.block(draw++, 0);
      pos = scv::m3(u01(b.x) * L, u01(b.y) * L, u01(b.z) * L);
    }
    scv::v4 rot = random_rotation();
    const int start = (int)A.size();
    for (size_t k = 0; k < types.size(); k++) {
      float radius, prev;
      int hyd;
      type_props(out, types[k], radius, hyd);
      if (k > 0) {
        type_props(out, types[k] - 1, prev, hyd);
        rot = scv::slerp(rot, random_rotation(), 0.3f);
        pos = scv::add(pos, scv::smul(-(prev + radius), scv::rotate(scv::up3(), rot)));
      }
      if (!push(types[k], pos, rot)) return false;
    }
    if (link) link_serially(A, start, (int)A.size(), bond_type);
  }
  return true;
}

bool build_setup(const char* config_json, const char* setup_json, HostSetup& out, std::string& err) {
  scj::Value cfg, setup;
  if (!config_json || !scj::parse(config_json, cfg, err)) { err = "config: " + (config_json ? err : std::string("null")); return false; }
  if (cfg.kind != scj::Value::Object) { err = "config: not an object"; return false; }
  if (setup_json && *setup_json) {
    if (!scj::parse(setup_json, setup, err)) { err = "setup: " + err; return false; }
  } else {
    setup.kind = scj::Value::Object;
  }
  sc_params& P = out.P;
  std::memset(&P, 0, sizeof(P));
  // defaults: spatial_cell_sim/config.json
  P.num_particles = cfg.get_int("numParticles", 65536);
  P.num_metabolic_particles = cfg.get_int("numMetabolicParticles", 65536);
  P.steps = cfg.get_int("steps", 401);
  P.sim_size = cfg.get_float("simSize", 400.0f);
  P.max_interaction_distance = cfg.get_float("maxInteractionDistance", 8.0f);
  P.max_diffusion_distance = cfg.get_float("maxDiffusionDistance", 10.0f);
  P.grid_cell_size = cfg.get_float("gridCellSize", 10.0f);
  P.movement_noise_scale = cfg.get_float("movementNoiseScale", 0.7f);
  P.rotation_noise_scale = cfg.get_float("rotationNoiseScale", 0.05f);
  P.metabolite_movement_noise_scale = cfg.get_float("metaboliteMovementNoiseScale", 1.5f);
  P.velocity_decay = cfg.get_float("velocityDecay", 0.7f);
  P.angular_velocity_decay = cfg.get_float("angularVelocityDecay", 0.7f);
  P.noise_coordination_steps = cfg.get_int("noiseCoordinationSteps", 16);
  P.relaxation_steps = cfg.get_int("relaxationSteps", 32);
  P.metabolite_diffusion_steps = cfg.get_int("metaboliteDiffusionSteps", 8);
  P.persist_every_nth_frame = cfg.get_int("persistEveryNthFrame", 1);
  P.report_every_nth_frame = cfg.get_int("reportEveryNthFrame", 10);
  std::string boundary = cfg.get_str("boundary", "clamp");
  if (boundary == "clamp") P.boundary = SC_BOUNDARY_CLAMP;
  else if (boundary == "periodic") P.boundary = SC_BOUNDARY_PERIODIC;
  else { err = "config: boundary must be \"clamp\" or \"periodic\""; return false; }
  P.propagate_state = cfg.get_bool("propagateState", false) ? 1 : 0;
  // Brownian noise model (extension): the reference's dimensionless scales,
  // or Stokes-Einstein coefficients of a sphere in a viscous fluid
  std::string noise = cfg.get_str("noiseModel", "reference");
  if (noise == "reference") {
    P.noise_model = SC_NOISE_REFERENCE;
  } else if (noise == "stokes_einstein") {
    P.noise_model = SC_NOISE_STOKES_EINSTEIN;
    const double T = cfg.get_double("temperature", 310.0);      // K
    const double eta = cfg.get_double("viscosity", 6.9e-4);     // Pa s (water near 310 K)
    const double dt = cfg.get_double("timestep", 0.1);          // ns
    if (!(T > 0.0) || !(eta > 0.0) || !(dt > 0.0)) { err = "config: temperature, viscosity and timestep must be positive"; return false; }
    const double kT = 1.380649e-23 * T;                        // J
    const double pi = 3.14159265358979323846;
    // 6 D(r) dt r = kT dt / (pi eta); 4 D_r(r) dt r^3 / E[theta^2 of a uniform rotation] with
    // E[theta^2] = pi^2 / 3 + 2; m^3 -> nm^3 is 1e27, ns -> s is 1e-9
    P.se_translation = (float)(kT * dt * 1e-9 / (pi * eta) * 1e27);
    P.se_rotation = (float)(kT * dt * 1e-9 / (2.0 * pi * eta) / (pi * pi / 3.0 + 2.0) * 1e27);
  } else {
    err = "config: noiseModel must be \"reference\" or \"stokes_einstein\"";
    return false;
  }
  uint64_t seed = (uint64_t)cfg.get_double("seed", 42.0);
  P.seed_lo = (uint32_t)seed;
  P.seed_hi = (uint32_t)(seed >> 32);
  if (!derive_grid(P, out.n_grid_cells, out.n_grid_bits, out.n_periodic, err)) { err = "config: " + err; return false; }
  if (P.grid_cell_size < P.max_interaction_distance)   // main.cu:153-158
    out.warnings.push_back("maxInteractionDistance exceeds gridCellSize; interactions beyond one cell are missed");
  if (P.grid_cell_size < P.max_diffusion_distance)
    out.warnings.push_back("maxDiffusionDistance exceeds gridCellSize; diffusion beyond one cell is missed");

  // ---- metabolite reactions (extension) ----
  if (const scj::Value* rx = setup.find("reactions")) {
    if (rx->size() > SC_MAX_REACTIONS) {
      err = "reactions: at most " + std::to_string(SC_MAX_REACTIONS) + " rows";
      return false;
    }
    for (size_t i = 0; i < rx->size(); i++) {
      const scj::Value& r = (*rx)[i];
      sc_reaction q;
      q.from = r.get_int("from", -1);
      q.to = r.get_int("to", -1);
      q.catalyst_type = r.get_int("catalystType", -1);
      q.rate = r.get_float("rate", 0.0f);
      q.catalyst_distance = r.get_float("catalystDistance", P.max_interaction_distance);
      const std::string at = "reactions[" + std::to_string(i) + "]: ";
      if (q.from < 0 || q.from >= SC_NUM_METABOLITES || q.to < 0 || q.to >= SC_NUM_METABOLITES || q.from == q.to) {
        err = at + "from and to must be distinct metabolite indices below " + std::to_string(SC_NUM_METABOLITES);
        return false;
      }
      if (!(q.rate >= 0.0f)) { err = at + "rate must be >= 0"; return false; }
      if (q.catalyst_type >= 0 && !(q.catalyst_distance > 0.0f && q.catalyst_distance <= P.grid_cell_size)) {
        err = at + "catalystDistance must be in (0, gridCellSize]";
        return false;
      }
      out.reactions.push_back(q);
    }
  }

  // ---- type table (chemistry.cuh:10-51) ----
  scj::Value types_doc, inter_doc;
  if (!load_doc(setup, "types", "types_file", types_doc, err)) return false;
  if (!load_doc(setup, "interactions", "interactions_file", inter_doc, err)) return false;
  scj::Value complex_doc;   // interactions.cuh:149-181 (gen/complexes.json)
  if (!load_doc(setup, "complexes", "complexes_file", complex_doc, err)) return false;
  int max_type = -1;
  for (size_t i = 0; i < types_doc.size(); i++) max_type = std::max(max_type, types_doc[i].get_int("type", 0));
  out.types.assign((size_t)(max_type + 1), sc_type_info{0.0f, 0});
  for (size_t i = 0; i < types_doc.size(); i++) {
    const scj::Value& m = types_doc[i];
    int t = m.get_int("type", 0);
    if (t < 0) { err = "types: negative type id"; return false; }
    out.types[t].radius = m.get_float("radius", 0.0f);
    out.types[t].hydrophobic = (int8_t)m.get_int("hydrophobic", 0);
  }

  // ---- interaction tables (interactions.cuh:17-145) ----
  std::vector<sc_interaction_info> rows;
  int max_id = 0;
  for (size_t i = 0; i < inter_doc.size(); i++) {
    rows.push_back(parse_interaction(inter_doc[i]));
    if (rows.back().id < 0) { err = "interactions: negative id"; return false; }
    if (rows.back().bond_probability > 0.0f && rows.back().group < 0) {
      err = "interactions: a row with bondProbability needs a group (its binding site)";
      return false;
    }
    max_id = std::max(max_id, rows.back().id);
  }
  std::sort(rows.begin(), rows.end(), [](const sc_interaction_info& a, const sc_interaction_info& b) { return a.id < b.id; });
  for (size_t i = 1; i < rows.size(); i++)
    if (rows[i].id == rows[i - 1].id) { err = "interactions: duplicate id " + std::to_string(rows[i].id); return false; }
  out.flat.assign((size_t)max_id + 1, sc_interaction_info());
  std::memset(out.flat.data(), 0, out.flat.size() * sizeof(sc_interaction_info));
  for (auto& r : rows) out.flat[r.id] = r;
  // partner map: per type (ascending), rows in ascending id order
  std::map<int, std::vector<sc_interaction_info>> by_type;
  int max_pt = max_id;
  for (auto& r : rows) {
    if (r.first_partner_type >= 0) { by_type[r.first_partner_type].push_back(r); max_pt = std::max(max_pt, r.first_partner_type); }
    if (r.second_partner_type >= 0) { by_type[r.second_partner_type].push_back(r); max_pt = std::max(max_pt, r.second_partner_type); }
  }
  max_pt = std::max(max_pt, max_type);
  out.partner_index.assign(2 * (size_t)(max_pt + 1), 0);
  out.partner.clear();
  for (auto& kv : by_type) {
    out.partner_index[2 * kv.first] = (int32_t)out.partner.size();
    for (auto& r : kv.second) out.partner.push_back(r);
    out.partner_index[2 * kv.first + 1] = (int32_t)out.partner.size();
  }

  // ---- populations ----
  out.run_init = setup.get_bool("init", true);
  const scj::Value* pops = setup.find("populations");
  const float L = P.sim_size;
  for (size_t pi = 0; pops && pi < pops->size(); pi++) {
    const scj::Value& d = (*pops)[pi];
    std::string arr = d.get_str("array", "plain");
    bool met = (arr == "metabolic");
    if (!met && arr != "plain") { err = "populations: array must be \"plain\" or \"metabolic\""; return false; }
    std::vector<sc_particle>& A = met ? out.metabolic : out.plain;
    int cap = met ? P.num_metabolic_particles : P.num_particles;
    int type = d.get_int("type", met ? SC_TYPE_METABOLIC : 0);
    int count = d.get_int("count", 0);
    if (d.has("count_fraction")) count = (int)((double)cap * d.get_double("count_fraction", 0.0));   // setup.cuh:29,236
    if (count < 0) { err = "populations: negative count"; return false; }
    std::string shape = d.get_str("shape", "uniform");
    if (shape == "sphere" || shape == "line" || shape == "wrapped_chain" || shape == "complex") {
      if (!place_shape(d, shape, pi, A, cap, out, complex_doc, err)) return false;
      continue;
    }
    if ((int64_t)A.size() + count > cap) { err = "populations exceed the capacity of the " + arr + " array"; return false; }
    float radius = 0.0f;
    int hyd = 0;
    bool known = type >= 0 && type < (int)out.types.size() && out.types[type].radius > 0.0f;
    if (known) { radius = out.types[type].radius; hyd = out.types[type].hydrophobic; }
    if (d.has("radius")) radius = d.get_float("radius", radius);
    if (d.has("hydrophobic")) hyd = d.get_int("hydrophobic", hyd);
    if (!(radius > 0.0f)) { err = "populations: type " + std::to_string(type) + " has no radius"; return false; }
    float center[3] = {0.5f * L, 0.5f * L, 0.5f * L};
    if (const scj::Value* c = d.find("center"))
      for (int k = 0; k < 3 && k < (int)c->size(); k++) center[k] = (*c)[k].as_float();
    Rng rng = {P.seed_lo, P.seed_hi, kSetupStream, (uint32_t)pi};
    int chain_len = std::max(1, d.get_int("chain_length", 1));
    float spacing = d.get_float("spacing", 2.0f * radius);
    int bond_type = d.get_int("bond_type", SC_INTERACTION_TYPE_CHAIN);
    float shell_r = d.get_float("shell_radius", 0.4f * L);
    float thickness = d.get_float("thickness", 0.0f);
    float ball_r = d.get_float("ball_radius", 0.25f * L);
    bool random_rot = d.get_bool("random_orientation", true);
    int start = (int)A.size();
    float chain_start[3] = {0, 0, 0}, chain_dir[3] = {0, 0, 1};
    for (int k = 0; k < count; k++) {
      sc_particle p = default_particle((int)A.size(), type, radius, hyd);
      scm::u32x4 b0 = rng.block((uint32_t)k, 0), b1 = rng.block((uint32_t)k, 1);
      float pos[3];
      if (shape == "uniform") {
        pos[0] = u01(b0.x) * L; pos[1] = u01(b0.y) * L; pos[2] = u01(b0.z) * L;
      } else if (shape == "shell" || shape == "ball") {
        float dir[3];
        unit_vector(u01(b0.x), u01(b0.y), dir);
        float r = (shape == "shell") ? shell_r + (u01(b0.z) - 0.5f) * thickness
                                     : ball_r * (float)std::cbrt((double)u01(b0.z));
        for (int q = 0; q < 3; q++) pos[q] = clampL(center[q] + r * dir[q], L);
      } else if (shape == "chains") {
        int m = k % chain_len;
        if (m == 0) {
          scm::u32x4 b2 = rng.block((uint32_t)k, 2);
          chain_start[0] = u01(b0.x) * L; chain_start[1] = u01(b0.y) * L; chain_start[2] = u01(b0.z) * L;
          unit_vector(u01(b2.x), u01(b2.y), chain_dir);
        }
        for (int q = 0; q < 3; q++) pos[q] = clampL(chain_start[q] + chain_dir[q] * (spacing * (float)m), L);
      } else {
        err = "populations: unknown shape '" + shape + "'";
        return false;
      }
      p.pos[0] = pos[0]; p.pos[1] = pos[1]; p.pos[2] = pos[2];
      if (random_rot) shoemake(u01(b1.x), u01(b1.y), u01(b1.z), p.rot);
      A.push_back(p);
    }
    if (shape == "chains" && chain_len > 1)
      for (int c0 = start; c0 < (int)A.size(); c0 += chain_len)
        link_serially(A, c0, std::min(c0 + chain_len, (int)A.size()), bond_type);
  }
  // metabolites.cuh seeding rule of setup/metabolite_setup.cuh:13-28 (slot 0 is skipped)
  if (const scj::Value* seeds = setup.find("metabolite_seed")) {
    for (size_t s = 0; s < seeds->size(); s++) {
      const scj::Value& sd = (*seeds)[s];
      int idx = sd.get_int("index", 0);
      float value = sd.get_float("value", 100.0f);
      double frac = sd.get_double("x_below_fraction", 0.2);
      if (idx < 0 || idx >= SC_NUM_METABOLITES) { err = "metabolite_seed: index out of range"; return false; }
      for (size_t i = 1; i < out.metabolic.size(); i++)
        if ((double)out.metabolic[i].pos[0] < (double)P.sim_size * frac) out.metabolic[i].metabolites[idx] = value;
    }
  }
  out.counts.n_active = (int32_t)out.plain.size();
  out.counts.n_active_metabolic = (int32_t)out.metabolic.size();
  out.counts.next_particle_id = (int32_t)out.plain.size();   // setup.cuh:223-225
  out.counts.step = 0;
  return true;
}

}  // namespace scs
