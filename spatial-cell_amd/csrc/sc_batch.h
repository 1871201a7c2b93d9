// Internal interface between the C ABI (sc_api.hip) and the serving loop
// (sc_server.cpp): a batch with reserved capacity whose instances join and
// leave between steps.
#pragma once
#include <cstdint>
#include <string>

#include "../../include/spatial_cell.h"
#include "sc_setup.h"

namespace scb {

struct ServerLimits {
  int cap_plain;        // plain slots of all resident instances together
  int cap_metabolic;    // metabolic slots
  int max_instances;    // resident instances
  uint32_t max_cells;   // cell keys of all resident instances (each aligned to its own size)
};

// Empty batch for instances that share the template's constants and tables.
sc_status server_sim_create(const scs::HostSetup& templ, const ServerLimits& lim, int device, sc_sim** out,
                            std::string& err);
// Instance constants must match the template (same rule as sc_batch_create).
bool server_compatible(const scs::HostSetup& templ, const scs::HostSetup& h, std::string& why);
// Instance `slot` (unused) joins: its particles are appended, its cell keys
// placed, its init sequence run on it alone; it starts at the current step.
// SC_ERR_BUDGET when its cells do not fit in the key space.
sc_status server_admit(sc_sim* sim, int slot, const scs::HostSetup& h);
// Kernel grids cover at least this many slots per array (rounded up to a
// power of two); every slot in use must be below it.
void server_set_launch(sc_sim* sim, size_t slots_plain, size_t slots_metabolic);
// Instance `slot` leaves: its slots are dropped by the next sort.
sc_status server_evict(sc_sim* sim, int slot);
const char* sim_error(const sc_sim* sim);
// Device bytes the batch allocates per plain / metabolic slot.
double bytes_per_slot(bool metabolic);

}  // namespace scb
