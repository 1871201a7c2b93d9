// Host-only entry points of the C ABI (no HIP calls): scene construction.
#include <string>

#include "../../include/spatial_cell.h"
#include "sc_setup.h"

struct sc_setup {
  scs::HostSetup h;
};

static thread_local std::string g_setup_error;

extern "C" {

sc_status sc_setup_create(const char* config_json, const char* setup_json, sc_setup** out) {
  if (!out) return SC_ERR_INVALID;
  *out = nullptr;
  sc_setup* s = new sc_setup();
  std::string err;
  if (!scs::build_setup(config_json, setup_json, s->h, err)) {
    g_setup_error = err;
    delete s;
    return SC_ERR_INVALID;
  }
  *out = s;
  return SC_OK;
}

const char* sc_setup_last_error(void) { return g_setup_error.c_str(); }

sc_status sc_setup_describe(const sc_setup* s, sc_params* params, sc_tables* tables, sc_counts* counts,
                            const sc_particle** plain, size_t* n_plain,
                            const sc_particle** metabolic, size_t* n_metabolic) {
  if (!s) return SC_ERR_INVALID;
  const scs::HostSetup& H = s->h;
  if (params) *params = H.P;
  if (tables) {
    tables->n_types = (int32_t)H.types.size(); tables->types = H.types.data();
    tables->n_flat = (int32_t)H.flat.size(); tables->flat = H.flat.data();
    tables->n_partner_index = (int32_t)H.partner_index.size(); tables->partner_index = H.partner_index.data();
    tables->n_partner = (int32_t)H.partner.size(); tables->partner = H.partner.data();
    tables->n_reactions = (int32_t)H.reactions.size(); tables->reactions = H.reactions.data();
  }
  if (counts) *counts = H.counts;
  if (plain) *plain = H.plain.data();
  if (n_plain) *n_plain = H.plain.size();
  if (metabolic) *metabolic = H.metabolic.data();
  if (n_metabolic) *n_metabolic = H.metabolic.size();
  return SC_OK;
}

sc_status sc_setup_destroy(sc_setup* s) {
  delete s;
  return SC_OK;
}

}  // extern "C"
