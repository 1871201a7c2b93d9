// Host-side construction of a simulation: config parsing, the type and
// interaction tables, and the initial particle arrays.
#pragma once
#include <string>
#include <vector>

#include "../../include/spatial_cell.h"

namespace scs {

struct HostSetup {
  sc_params P;
  int n_grid_cells = 0;       // power of two per axis (config.cuh:66-74)
  int n_grid_bits = 0;
  int n_periodic = 0;         // L / cell for the periodic boundary
  std::vector<sc_type_info> types;
  std::vector<sc_interaction_info> flat;
  std::vector<sc_interaction_info> partner;
  std::vector<int32_t> partner_index;
  std::vector<sc_reaction> reactions;
  std::vector<sc_particle> plain;       // slots [0, n_active)
  std::vector<sc_particle> metabolic;   // slots [0, n_active_metabolic)
  sc_counts counts;
  bool run_init = true;
  std::vector<std::string> warnings;
};

// Parses config_json (keys of spatial_cell_sim/config.json + extensions) and
// setup_json; generates the populations.  Returns false with err set on a
// malformed or inconsistent document.
bool build_setup(const char* config_json, const char* setup_json, HostSetup& out, std::string& err);

// Derived grid values of config.cuh:66-74 plus validation; used by build_setup
// and by restore paths.
bool derive_grid(const sc_params& P, int& n_cells, int& bits, int& n_periodic, std::string& err);

// A particle record with the field defaults of the reference constructor
// (types.cuh:185-214).
sc_particle default_particle(int id, int type, float radius, int hydrophobic);

}  // namespace scs
