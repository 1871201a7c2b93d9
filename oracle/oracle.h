/* ORACLE -- test infrastructure only.  Serial host-C++ restatement of the
 * reference step (spatial_cell_sim/src/main.cu:347-700 and the kernels it
 * launches).  Used by tests/ as the parity checker and by bench.py as the CPU
 * baseline; the product library never links it.
 *
 * Every part is pinned by the tests under tests/test_cpu_*.py: closed forms,
 * brute force, textbook equivalents and known answers; the bond-geometry
 * terms of relax (dynamics.cuh:585-770) and coordinateNoise
 * (dynamics.cuh:288-402) by special cases whose answers follow from the
 * geometry.  removeInterfering, move's debugVector reset and the periodic wrap
 * of applyVelocities are pinned by tests/test_cpu_oracle_pins.py, which also
 * shows that planted mistakes (orc_plant, ORACLE_PLANTS build) fail them. */
#ifndef SC_ORACLE_H
#define SC_ORACLE_H
#include "../include/spatial_cell.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

/* Copies params, tables and the state (slots [0, n) of each array; the rest of
 * each capacity is zero, as after the reference's host memset). */
orc_sim* orc_create(const sc_params* params, const sc_tables* tables, const sc_counts* counts,
                    const sc_particle* plain, size_t n_plain,
                    const sc_particle* metabolic, size_t n_metabolic);
void orc_destroy(orc_sim* s);
/* One full step of main.cu:347-649 per iteration (no frame persistence). */
void orc_step(orc_sim* s, int n);
/* One stage, same ids and arguments as sc_run_stage. Returns 0 or -1. */
int orc_run_stage(orc_sim* s, int stage, int arg);
void orc_snapshot(const orc_sim* s, sc_counts* counts, sc_particle* plain, size_t plain_cap,
                  sc_particle* metabolic, size_t metabolic_cap);
/* Same layouts as sc_debug_export. Returns bytes needed / written. */
size_t orc_debug_export(const orc_sim* s, int what, int population, void* dst, size_t cap);
/* frames.dat frame of the current state (reduceParticles + writeFrame). */
size_t orc_frame_size(const orc_sim* s);
size_t orc_emit_frame(const orc_sim* s, void* dst, size_t cap);

/* Test hooks: scalar primitives (same ids as sc_eval_primitive), raw
 * Philox4x32-10 blocks, and the vector/quaternion helpers of math.cuh. */
void orc_eval_primitive(int fn, const float* x, float* y, size_t n);
void orc_philox(const uint32_t* ctr4, const uint32_t* key2, uint32_t* out4, size_t n);
void orc_eval_math(int fn, const float* in9, float* out4, size_t n);

#ifdef __cplusplus
}
#endif
#endif
