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

// Copying a retiring instance's final state out without holding the batch:
// its slots are gathered on the device into a stage's device half (stream
// order keeps the next steps behind the gather), then copied to the stage's
// pinned host half on the copy stream while the batch steps on; `done`
// (a hipEvent_t) fires when the host half is complete.
struct RetireStage {
  void* dev = nullptr;
  void* host = nullptr;
  size_t bytes = 0;
  void* done = nullptr;
};
struct Retired {
  size_t cap[2] = {0, 0};   // the instance's capacities: rows of its gathered fields
  int step = 0;             // its step count when it retired
};
size_t server_retire_bytes(size_t cap_plain, size_t cap_metabolic);
sc_status server_stage_alloc(sc_sim* sim, size_t bytes, RetireStage& st);
void server_stage_free(sc_sim* sim, RetireStage& st);
// Enqueues the gather and the copy of instance `slot` (before server_evict).
sc_status server_retire_begin(sc_sim* sim, int slot, const RetireStage& st, Retired& r);
bool server_retire_ready(const RetireStage& st);
// Waits for the batch's stream and its copy stream (after a failed retire:
// no copy into a stage is left in flight).
void server_drain(sc_sim* sim);
void server_retire_counts(const RetireStage& st, const Retired& r, sc_counts* out);
// Records of the host half into plain / metabolic (zero records past the
// instance's slots up to the caps).
void server_retire_decode(const RetireStage& st, const Retired& r, sc_particle* plain, size_t plain_cap,
                          sc_particle* metabolic, size_t metabolic_cap);
// Device bytes the batch allocates per plain / metabolic slot.
double bytes_per_slot(bool metabolic);

}  // namespace scb
