/*
 * spatial_cell.h -- C ABI of the MI355X (gfx950) per-step particle pipeline of
 * spatial_cell_sim.  Plain C types only: a caller binds these entry points
 * through ctypes/cgo/JNI/N-API exactly like any other C library.
 *
 * The reference has no plugin or FFI surface: main() owns the whole
 * simulation (spatial_cell_sim/src/main.cu:142-721).  The seams replaced here
 * are the init block (main.cu:145-338), the body of the step loop
 * (main.cu:347-700) and frame persistence (main.cu:95-126,
 * spatial_cell_sim/src/storage.cuh:22-39).  Each entry point cites the part it
 * replaces.
 *
 * Conventions: every function returns an sc_status; on failure the message is
 * available from sc_last_error().  A handle is not thread-safe: calls on one
 * handle must be serialised by the caller.  The library never exits the
 * process (the reference calls exit(EXIT_FAILURE), memory.cuh:12-16).
 */
#ifndef SPATIAL_CELL_H
#define SPATIAL_CELL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sc_status {
  SC_OK = 0,
  SC_ERR_INVALID = 1,  /* bad argument, config or setup */
  SC_ERR_OOM = 2,      /* device or host allocation failed */
  SC_ERR_DEVICE = 3,   /* HIP runtime error */
  SC_ERR_BUDGET = 4,   /* capacity exceeded */
  SC_ERR_STATE = 5     /* call not valid in the handle's current state */
} sc_status;

/* Compile-time caps, identical to spatial_cell_sim/src/constants.cuh:17-25. */
#define SC_MAX_BONDS 8
#define SC_NUM_METABOLITES 10
#define SC_REDUCED_NUM_METABOLITES 4
#define SC_FLAG_ACTIVE 0x0001
#define SC_TYPE_LIPID 0
#define SC_TYPE_METABOLIC 1000
#define SC_INTERACTION_TYPE_CHAIN 1
#define SC_GROUP_FORWARD 10
#define SC_GROUP_BACKWARD 11
#define SC_HYDROPHILIC 0
#define SC_HYDROPHOBIC 1
#define SC_POLAR 2

#define SC_BOUNDARY_CLAMP 0     /* reference: positions clamped to [0, L], no wrap */
#define SC_BOUNDARY_PERIODIC 1  /* extension: periodic cube, wrapped stencil */

/* Runtime parameters.  The first 17 fields are the keys of
 * spatial_cell_sim/config.json, parsed as in config.cuh:46-63; the rest are
 * extensions without a reference counterpart. */
typedef struct sc_params {
  int32_t num_particles;            /* capacity of the plain array */
  int32_t num_metabolic_particles;  /* capacity of the metabolic array */
  int32_t steps;
  float sim_size;
  float max_interaction_distance;
  float max_diffusion_distance;
  float grid_cell_size;
  float movement_noise_scale;
  float rotation_noise_scale;
  float metabolite_movement_noise_scale;
  float velocity_decay;
  float angular_velocity_decay;
  int32_t noise_coordination_steps;
  int32_t relaxation_steps;
  int32_t metabolite_diffusion_steps;
  int32_t persist_every_nth_frame;
  int32_t report_every_nth_frame;
  /* extensions */
  int32_t boundary;                 /* SC_BOUNDARY_* */
  uint32_t seed_lo, seed_hi;        /* Philox key */
  int32_t propagate_state;          /* 1: propagateState runs after polymerize in each step */
  /* Brownian noise (dynamics.cuh:156-215): SC_NOISE_REFERENCE scales by
   * movementNoiseScale / max(r / 2.5, 1) and rotationNoiseScale / (r / 2.5);
   * SC_NOISE_STOKES_EINSTEIN uses the Stokes-Einstein coefficients of a
   * sphere of radius r (nm) in a fluid of viscosity eta at temperature T
   * over a timestep dt: displacement scale sqrt(se_translation / r), so each
   * axis has variance 2 D dt with D = kT / (6 pi eta r); rotation fraction
   * sqrt(se_rotation / r^3), so the mean squared rotation angle is 4 D_r dt
   * with D_r = kT / (8 pi eta r^3).  Both coefficients in nm^3, computed from
   * the config keys "temperature" (K), "viscosity" (Pa s), "timestep" (ns). */
  int32_t noise_model;
  float se_translation;
  float se_rotation;
} sc_params;

#define SC_NOISE_REFERENCE 0
#define SC_NOISE_STOKES_EINSTEIN 1

/* Device/host counters (reference: nActiveParticles, nActiveMetabolicParticles,
 * nextParticleId, main.cu:244-253; step is the loop index of main.cu:347). */
typedef struct sc_counts {
  int32_t n_active;
  int32_t n_active_metabolic;
  int32_t next_particle_id;
  int32_t step;
} sc_counts;

/* Full FP32 state of one particle (reference struct Particle /
 * MetabolicParticle, types.cuh:167-236), as a flat 4-byte-field record. */
typedef struct sc_particle {
  int32_t id, type, flags, state;
  int32_t last_state_change_step, hydrophobic, n_bonds, scheduled_polymerization;
  float radius;
  float pos[3];
  float rot[4];
  float vel[3];
  float pos_noise[3];
  float ang_vel[4];
  float ang_noise[4];
  float debug[4];
  int32_t bonds[SC_MAX_BONDS][3];   /* {type, group, partnerId}, types.cuh:161-165 */
  float metabolites[SC_NUM_METABOLITES];
} sc_particle;

/* Type table entry (reference MinimalParticleTypeInfo, types.cuh:10-26). */
typedef struct sc_type_info {
  float radius;
  int32_t hydrophobic;
} sc_type_info;

/* Interaction table row (reference ParticleInteractionInfo, types.cuh:49-120).
 * States are stored widened from the reference's int8. */
typedef struct sc_interaction_info {
  int32_t id, group;
  int32_t first_partner_type, first_partner_state;
  int32_t second_partner_type, second_partner_state;
  int32_t propagate_state, set_state, wait_for_state, wait_for_alignment;
  int32_t only_via_transition, transition_to, polymerize, break_after_transition;
  float relative_orientation[4];
  float orientation_alignment_threshold;
  float relative_position[3];
  float position_alignment_threshold;
  /* Extension (no reference counterpart): a row with bond_probability > 0
   * forms its bond by site contact instead of the alignment thresholds of
   * interactions.cuh:319-333: the second partner's centre within
   * contact_distance of the first partner's site (its position plus the
   * row's relative position in its frame), a Philox draw below
   * bond_probability keyed by the two ids, the row and the step, both
   * partners free in the row's group and choosing each other (the lowest
   * (partner id, row id) among a group's candidates); the bond is then added
   * to both. */
  float bond_probability;
  float contact_distance;
} sc_interaction_info;

/* Metabolite reaction (extension: the reference only diffuses metabolites,
 * metabolites.cuh:352-431).  Once per step, after the diffusion passes, each
 * active metabolic particle converts the fraction min(rate * n, 1) of
 * metabolite `from` into metabolite `to`, n = 1 for an uncatalysed reaction,
 * else the number of active plain particles of catalyst_type whose centres
 * lie within catalyst_distance (at most gridCellSize) of the particle's.
 * Reactions apply in table order, each to the values the last one left. */
#define SC_MAX_REACTIONS 16

typedef struct sc_reaction {
  int32_t from, to;           /* metabolite indices, 0 .. SC_NUM_METABOLITES - 1 */
  int32_t catalyst_type;      /* plain particle type, or -1: uncatalysed */
  float rate;                 /* fraction of m[from] per step (per catalyst) */
  float catalyst_distance;    /* nm */
} sc_reaction;

/* Flattened tables as built by the reference loaders
 * (chemistry.cuh:33-51, interactions.cuh:59-145), and the reaction list. */
typedef struct sc_tables {
  int32_t n_types;                          /* types[t] for t < n_types */
  const sc_type_info* types;
  int32_t n_flat;                           /* flat[id] for id < n_flat */
  const sc_interaction_info* flat;
  int32_t n_partner_index;                  /* partner_index[2*type], [2*type+1] */
  const int32_t* partner_index;
  int32_t n_partner;                        /* partner-mapped rows */
  const sc_interaction_info* partner;
  int32_t n_reactions;                      /* setup document key "reactions" */
  const sc_reaction* reactions;
} sc_tables;

/* Frame records (reference ReducedParticle / ReducedMetabolicParticle,
 * types.cuh:238-294): 36 and 44 bytes, fields little-endian, halves IEEE
 * binary16 rounded to nearest even. */
#define SC_REDUCED_PARTICLE_BYTES 36
#define SC_REDUCED_METABOLIC_BYTES 44
#define SC_FRAME_HEADER_BYTES 12

typedef struct sc_sim sc_sim;

/* Pipeline stages, in the order of one reference step (main.cu:347-649).
 * Used by sc_run_stage and sc_profile_steps.  The multi-pass stages (relax,
 * metabolic relax, partners, diffusion) build their per-step neighbour lists
 * at pass arg = 0: a sequence of passes must start with arg 0, as sc_step's
 * does. */
enum {
  SC_STAGE_SORT = 0,               /* updateGridAndSort x2, main.cu:350-367 */
  SC_STAGE_MOVE = 1,               /* move + brownian (both arrays), main.cu:375-421 */
  SC_STAGE_COORDINATE = 2,         /* one coordinateNoise pass, arg = j, main.cu:426-441 */
  SC_STAGE_APPLY_NOISE = 3,        /* applyNoise x2, main.cu:443-455 */
  SC_STAGE_RELAX = 4,              /* one relax pass, arg = j, main.cu:487-496 */
  SC_STAGE_RELAX_METABOLIC = 5,    /* relaxMetabolicParticles, arg = j, main.cu:510-523 */
  SC_STAGE_RELAX_PARTNERS = 6,     /* relaxMetabolicParticlePartners, arg = j, main.cu:526-537 */
  SC_STAGE_APPLY_VELOCITIES = 7,   /* applyVelocities x2, main.cu:548-561 */
  SC_STAGE_COMPLEXIFY = 8,         /* complexify, main.cu:567-579 (arg = step) */
  SC_STAGE_TRANSITION = 9,         /* transitionInteractions, main.cu:582-598 */
  SC_STAGE_POLYMERIZE = 10,        /* polymerize, main.cu:600-616 */
  SC_STAGE_DIFFUSE = 11,           /* one diffuseMetabolites pass, arg = j, main.cu:637-649 */
  SC_STAGE_REMOVE_INTERFERING = 12,/* init-time overlap removal, main.cu:301-312 */
  SC_STAGE_END_STEP = 13,          /* step counter += 1 */
  SC_STAGE_PROPAGATE_STATE = 14,   /* propagateState, interactions.cuh:743-822 (commented out at
                                      main.cu:618-631; in the step after polymerize when the config
                                      sets "propagateState": true) */
  SC_STAGE_REACT = 15,             /* metabolite reactions (sc_reaction; no reference counterpart),
                                      after the diffusion passes when the setup lists reactions */
  SC_NUM_STAGES = 16
};

/* Exports for parity checks (sc_debug_export 'what'). */
enum {
  SC_EXPORT_KEYS = 0,        /* u32 per slot: reference cell key of each sorted slot (grid.cuh:13-44) */
  SC_EXPORT_PERMUTATION = 1, /* i32 per slot: pre-sort slot of each sorted slot */
  SC_EXPORT_RANGES = 2,      /* u32 triples {key, start, end} for every non-empty cell, key order */
  SC_EXPORT_NEIGHBORS = 3,   /* i32 pairs {slot, neighbour slot} passing the box test, enumeration order */
  SC_EXPORT_LIST_COUNTS = 4, /* u32 pairs per active slot: entry counts of the step's two pair lists
                                (plain: relax, lipid back-reaction; metabolic: relax, diffusion);
                                0xffff = the particle walks the stencil in every pass instead */
  SC_EXPORT_FUSED = 5        /* u32 per active slot: 1 if the last step relaxed it in the fused kernel */
};

/* Host-only scene construction (no device needed): parses the same documents
 * as sc_create and generates the initial particle arrays.  Used to inspect a
 * scene or to hand it to another engine.  Pointers returned by
 * sc_setup_describe stay valid until sc_setup_destroy. */
typedef struct sc_setup sc_setup;
sc_status sc_setup_create(const char* config_json, const char* setup_json, sc_setup** out);
sc_status sc_setup_describe(const sc_setup* setup, sc_params* params, sc_tables* tables, sc_counts* counts,
                            const sc_particle** plain, size_t* n_plain,
                            const sc_particle** metabolic, size_t* n_metabolic);
sc_status sc_setup_destroy(sc_setup* setup);
/* Message of the last failed sc_setup_create on this thread. */
const char* sc_setup_last_error(void);

/* Create a simulation.  config_json: the keys of spatial_cell_sim/config.json
 * plus optional extensions ("boundary", "seed").  setup_json: particle types,
 * interaction table and the populations to generate (replaces the hard-coded
 * scene of setup/setup.cuh and the loaders, main.cu:160-175,262-265).  Runs the
 * init sequence of main.cu:283-338 (sort, overlap removal, complexify at step
 * -1) unless the setup sets "init": false.  device: HIP device ordinal. */
sc_status sc_create(const char* config_json, const char* setup_json, int device, sc_sim** out);
sc_status sc_destroy(sc_sim* sim);
/* Message of the last failed call on this handle; with sim == NULL, of the last
 * failed sc_create, sc_batch_create, sc_bin or sc_bin_noalloc on this thread. */
const char* sc_last_error(const sc_sim* sim);

/* ---- Batches: many independent cells in one launch chain ------------------
 * No reference counterpart: the reference runs one cell per process
 * (main.cu:145-700).  A batch steps n instances with one chain of kernel
 * launches; each instance evolves exactly as a single cell created from the
 * same two documents would (same slots in the same order, same Philox stream:
 * seed, particle id and the instance's own step).  Instances must agree on
 * gridCellSize, maxInteractionDistance, maxDiffusionDistance, boundary, the
 * noise and decay constants, the sub-iteration counts, the type and interaction
 * tables and "init"; they may differ in simSize, seed, capacities and
 * populations.  sc_step, sc_run_stage, sc_synchronize, sc_time_steps,
 * sc_profile_steps and sc_destroy take a batch handle; sc_snapshot, sc_restore,
 * the frame calls and sc_debug_export refuse one (use the calls below). */
sc_status sc_batch_create(int n, const char* const* config_json, const char* const* setup_json, int device,
                          sc_sim** out);
/* Number of instances (1 for a single-cell handle). */
sc_status sc_batch_size(const sc_sim* sim, int* n);
/* Counts of one instance: its plain / metabolic slots, next id and step. */
sc_status sc_batch_counts(sc_sim* sim, int instance, sc_counts* out);
/* Records of one instance's slots in slot order, as sc_snapshot of a single
 * cell would give them; records past its slot count are zero. */
sc_status sc_batch_snapshot(sc_sim* sim, int instance, sc_counts* counts, sc_particle* plain, size_t plain_cap,
                            sc_particle* metabolic, size_t metabolic_cap);
/* frames.dat header and frame of one instance (storage.cuh:22-39). */
sc_status sc_batch_frame_header(const sc_sim* sim, int instance, void* dst12);
sc_status sc_batch_frame_size(const sc_sim* sim, int instance, size_t* bytes);
sc_status sc_batch_emit_frame(sc_sim* sim, int instance, void* dst, size_t cap, size_t* written);

/* ---- Serving loop ------------------------------------------------------------
 * No reference counterpart.  Requests (one cell: config + setup documents and
 * a number of steps) arrive at any time; a scheduler thread admits them, first
 * come first served, into one running batch while the reserved slots, instance
 * slots and cell keys allow, steps the batch, and retires each request after
 * its steps: its final state is kept for sc_server_result and its slots leave
 * the batch at the next sort.  A request's result equals running its documents
 * alone for its steps.  server_json: {"max_particles", "max_metabolic_particles"
 * (slots), or "hbm_budget_bytes" [+ "metabolic_fraction", default 0.55] to
 * derive them; "max_instances" (default 256), "max_cells" (cell keys of all
 * resident cells, default 2^28), "max_retire_stage_bytes" (largest staging
 * buffer for a finished cell's final state, default no limit: a cell that
 * needs more fails, leaves the batch and frees its slot)}.  The template documents fix the constants
 * every request must share (as for sc_batch_create); their populations are
 * ignored. */
typedef struct sc_server sc_server;
enum { SC_REQ_QUEUED = 0, SC_REQ_RUNNING = 1, SC_REQ_DONE = 2, SC_REQ_FAILED = 3 };
typedef struct {
  int64_t id;
  int32_t state;          /* SC_REQ_* */
  int32_t steps;          /* requested */
  int32_t steps_done;
  int32_t pad;
  double t_submit, t_admit, t_done;   /* seconds since sc_server_start; -1 until it happens */
  sc_counts counts;       /* final counts once done */
} sc_request_info;
sc_status sc_server_start(const char* config_json, const char* setup_json, const char* server_json, int device,
                          sc_server** out);
/* SC_ERR_BUDGET: larger than the server's capacity. */
sc_status sc_server_submit(sc_server* srv, const char* config_json, const char* setup_json, int steps, int64_t* id);
sc_status sc_server_poll(sc_server* srv, int64_t id, sc_request_info* info);
/* Blocks until the request is done (SC_OK), failed (SC_ERR_DEVICE) or
 * timeout_s passed (SC_ERR_BUDGET). */
sc_status sc_server_wait(sc_server* srv, int64_t id, double timeout_s);
/* Final records of a done request (slot order, as sc_snapshot of the cell). */
sc_status sc_server_result(sc_server* srv, int64_t id, sc_particle* plain, size_t plain_cap,
                           sc_particle* metabolic, size_t metabolic_cap);
/* Forgets a done or failed request. */
sc_status sc_server_release(sc_server* srv, int64_t id);
/* Stops the scheduler (requests not done are abandoned) and frees the GPU state. */
sc_status sc_server_stop(sc_server* srv);
/* Last error of srv; with srv == NULL, of the last failed sc_server_start. */
const char* sc_server_last_error(sc_server* srv);
/* Where the scheduler thread's time went since sc_server_start: admitting
 * (upload and init of new cells), stepping (one step of the batch and the
 * wait for it), retiring (copying final states out), idle (no cell
 * resident); batch steps run and the particle slots they stepped (the sum
 * over steps of the resident cells' capacities).  No reference counterpart. */
typedef struct {
  double t_admit, t_step, t_retire, t_idle;   /* seconds */
  int64_t steps, admitted, retired;
  double slot_steps;
} sc_server_stats;
sc_status sc_server_get_stats(sc_server* srv, sc_server_stats* out);

/* ---- Binning of caller-owned arrays -----------------------------------------
 * updateGridAndSort (grid.cuh:111-139) as a routine over device arrays the
 * caller owns, with the kernels of the simulation's sort stage: the cell key
 * of every slot (getGridIdx / makeIdx / updateIndices, grid.cuh:13-44; an
 * inactive slot's key is 0xFFFFFFFF), the stable sort of the keys
 * (grid.cuh:76-108) and the cell table (updateGridRanges, grid.cuh:46-73) as
 * CSR starts.  The reference sorts the particle structs themselves
 * (grid.cuh:101-102); here perm_out[i] = the pre-sort slot of sorted slot i,
 * so the caller gathers whatever fields it keeps.
 *   pos:  n records {x, y, z, w} of float (w is not read), 16-byte aligned;
 *   hdr:  n records {id, type, flags, hydrophobic} of int32, 16-byte aligned
 *         (only flags bit 0, SC_FLAG_ACTIVE, is read);
 *   keys_out: n sorted keys; perm_out: n int32;
 *   cell_start: cells + 1 u32 with cells = nGridCells^3 (config.cuh:66-74,
 *         nGridCells = the power of two >= ceil(sim_size / grid_cell_size));
 *         the slots of cell k are [cell_start[k], cell_start[k + 1]), i.e. the
 *         reference's gridRanges[2k] and [2k + 1] (an empty cell: equal).
 * boundary: SC_BOUNDARY_* (periodic: the last cell along an axis is
 * L / cell - 1).  All arrays are device memory of `device`; stream is a
 * hipStream_t of that device, or NULL for the null stream.  Allocates its
 * workspace (about 16 B per slot plus cells / 4 B) and returns when the
 * results are complete.  Errors: sc_last_error(NULL). */
sc_status sc_bin(int device, const float* pos, const int32_t* hdr, int n, float sim_size, float grid_cell_size,
                 int boundary, uint32_t* keys_out, int32_t* perm_out, uint32_t* cell_start, void* stream);

/* sc_bin without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as sc_bin, but the workspace is a fixed array in the
 * library's device memory (about 21 MB, held by each GPU the library runs on
 * until the process ends), so the call allocates nothing, only enqueues its
 * kernels on `stream` and returns: the results are complete when the stream
 * reaches that point.  Limits: n <= SC_BIN_NOALLOC_MAX_SLOTS and at most
 * SC_BIN_NOALLOC_MAX_CELLS cells (a grid of up to 256 per axis); beyond them
 * it returns SC_ERR_BUDGET.  All calls on one GPU share its array: calls
 * that overlap on one GPU (other streams, other threads) race, so serialise
 * them (one stream, or wait for the previous call's stream). */
#define SC_BIN_NOALLOC_MAX_SLOTS (1 << 20)
#define SC_BIN_NOALLOC_MAX_CELLS (1 << 24)
sc_status sc_bin_noalloc(int device, const float* pos, const int32_t* hdr, int n, float sim_size,
                         float grid_cell_size, int boundary, uint32_t* keys_out, int32_t* perm_out,
                         uint32_t* cell_start, void* stream);

/* Advance n steps (one step = main.cu:347-649; frame persistence is separate). */
sc_status sc_step(sc_sim* sim, int n);
/* Run one stage (SC_STAGE_*) on the current state. */
sc_status sc_run_stage(sc_sim* sim, int stage, int arg);
sc_status sc_synchronize(sc_sim* sim);
sc_status sc_counts_get(sc_sim* sim, sc_counts* out);
sc_status sc_get_params(const sc_sim* sim, sc_params* out);
/* Pointers stay valid for the handle's lifetime. */
sc_status sc_get_tables(const sc_sim* sim, sc_tables* out);

/* Full FP32 state, slots [0, capacity) of each array (no reference
 * counterpart: the reference cannot checkpoint, see frames.dat). */
sc_status sc_snapshot(sc_sim* sim, sc_counts* counts,
                      sc_particle* plain, size_t plain_cap,
                      sc_particle* metabolic, size_t metabolic_cap);
sc_status sc_restore(sc_sim* sim, const sc_counts* counts,
                     const sc_particle* plain, size_t n_plain,
                     const sc_particle* metabolic, size_t n_metabolic);

/* frames.dat: 12-byte header (storage.cuh:22-27) and one frame
 * (reduceParticles + writeFrame, dynamics.cuh:824-844, storage.cuh:29-39). */
sc_status sc_frame_header(const sc_sim* sim, void* dst12);
sc_status sc_frame_size(const sc_sim* sim, size_t* bytes);
sc_status sc_emit_frame(sc_sim* sim, void* dst, size_t cap, size_t* written);
/* The same frame without waiting for it (the reference's asynchronous
 * writer, main.cu:95-126): the reduce is queued behind the steps already
 * queued, and the copy to dst runs on a copy stream beside later steps.
 * dst is complete after the next sc_frame_wait on this handle (or
 * sc_synchronize); the caller must not touch it before.  With dst from
 * sc_host_alloc (pinned) the copy is asynchronous; with other memory it
 * still works but the copy blocks.  The device frame buffer is reused: a
 * later frame's reduce waits for this copy. */
sc_status sc_emit_frame_async(sc_sim* sim, void* dst, size_t cap, size_t* written);
sc_status sc_frame_wait(sc_sim* sim);
/* Pinned host memory for frame buffers (hipHostMalloc / hipHostFree). */
sc_status sc_host_alloc(size_t bytes, void** out);
sc_status sc_host_free(void* p);

/* Parity exports (SC_EXPORT_*) for population 0 (plain) or 1 (metabolic).
 * With dst == NULL only *written (bytes needed) is set. */
sc_status sc_debug_export(sc_sim* sim, int what, int population,
                          void* dst, size_t cap, size_t* written);

/* Time n steps stage by stage with HIP events on the library's stream;
 * stage_ms[s] receives the average milliseconds per step spent in stage s
 * (n_stages <= SC_NUM_STAGES). */
sc_status sc_profile_steps(sc_sim* sim, int n, double* stage_ms, int n_stages);
/* Time n steps as one region with HIP events; *ms = average per step. */
sc_status sc_time_steps(sc_sim* sim, int n, double* ms);

/* Kernels sc_time_kernel can time on their own. */
enum {
  SC_KERNEL_RELAX_PASS = 0,      /* one later relax pass, bonded and unbonded particles (main.cu:487-496) */
  SC_KERNEL_METABOLIC_PASS = 1,  /* one later metabolic relax pass (main.cu:510-523) */
  SC_KERNEL_DIFFUSE_PASS = 2,    /* one later diffusion pass (main.cu:637-649) */
  SC_KERNEL_RELAX_FUSED = 3      /* the whole relaxation (main.cu:468-565) of the last step's fused
                                    components, written to the next buffers */
};
/* Launch one pass kernel reps + 1 times back to back on the library's
 * stream; each of the last reps records the start and end of its dispatch in
 * two HIP events of its own (hipExtLaunchKernelGGL); *ms = the mean of their
 * elapsed times (the kernel's time on the device).  The passes read
 * the current buffers and write the next ones without swapping, so the
 * simulation state is unchanged.  Needs the step's neighbour lists: call
 * after sc_step.  No reference counterpart (measurement). */
sc_status sc_time_kernel(sc_sim* sim, int kernel, int reps, double* ms);
/* Run `steps` steps (the simulation advances), launched kernel by kernel on
 * the library's stream; every launch of the kernel within them records the
 * start and end of its dispatch in HIP events of its own
 * (hipExtLaunchKernelGGL); *ms = the mean of their elapsed times: the
 * kernel's time on the device among the step's other kernels.  Only
 * SC_KERNEL_RELAX_PASS (all relaxationSteps launches of each step); needs
 * option "concurrent" 0.  No reference counterpart (measurement). */
sc_status sc_time_step_kernel(sc_sim* sim, int kernel, int steps, double* ms);
/* Autotuner (no reference counterpart; the reference's only tuning is the
 * hand-patched CUB policy of grid.cuh:86-90).  Times `steps` steps (after
 * one untimed step that captures the step graph) for each candidate setting
 * of the scheduling options whose results are bit-identical -- "relax_split"
 * with "coll_gather" and "pass_waves", then "xcd_run", then "sort_items",
 * then "fuse_bonds" where bonds can exist, each with the best of the earlier
 * stages -- and
 * keeps the fastest.  Every candidate starts from the same device-side checkpoint of
 * the whole state, restored at the end: the simulation does not advance.
 * Call it after a few steps (the list pools have grown by then).  report
 * (may be NULL) receives a JSON line: the choice and every candidate's
 * milliseconds per step.  Needs device memory for a copy of the state. */
sc_status sc_autotune(sc_sim* sim, int steps, char* report, size_t report_cap);

/* Evaluates one scalar primitive elementwise (0 sin, 1 cos, 2 acos, 3 log,
 * 4 log2, 5 exp2, 6 pow(x, 0.2), 7 sqrt, 8 float->binary16 bits, 9 Box-Muller
 * normal from the bits of x, 10 Philox4x32-10 word mix, 11 Shoemake rotation
 * of +y) on the device (on_device = 1) or with the host build of the same
 * code; both must agree bit for bit.  No reference counterpart: it checks that
 * host and device compilers round identically. */
sc_status sc_eval_primitive(int device, int fn, const float* x, float* y, size_t n, int on_device);
/* Options: "graph" (1 = replay each step as a captured hipGraph, default 1);
 * "xcd_run" (neighbour kernels: consecutive tiles of 256 sorted slots that one
 * XCD takes in turn, 0 = dispatch order (default); results do not depend on it);
 * "concurrent" (0 = one chain of launches, default; 1 = launches that touch
 * one population only run beside the other population's, as parallel
 * branches of the step's graph or on a second stream; 2 = also the bonded
 * relax pass beside the unbonded one; results do not depend on it);
 * "fuse" (1 = the connected components of the step's contact graph that fit
 * one wave, and have no bonded member, run all relaxation passes in one
 * launch with their velocities in LDS; 0 = every pass a launch of its own
 * over all particles (default); results do not depend on it);
 * "relax_split" (1 = every relax pass as two kernels run side by side: the
 * general one for particles with polar alignment entries, stencil walkers and
 * bonded particles whose bond terms the pass kernel computes ("fuse_bonds"),
 * and a lean collision-only one for the rest, bonded collision-only particles
 * included, which runs more waves per SIMD; 0 = one kernel for all
 * (default); results do not depend on it);
 * "pass_waves" (3 or 4: waves per SIMD the general relax pass kernel's
 * registers are sized for; 4 keeps a few values that live across its list
 * walk in scratch memory; default 3; results do not depend on it);
 * "coll_gather" (4 or 8: neighbours the collision-only relax kernel gathers
 * per memory round trip; 4 needs fewer registers and runs more waves per
 * SIMD; default 8; results do not depend on it);
 * "sort_items" (4, 8 or 16: keys per thread of a onesweep radix-sort tile,
 * tiles of 1024, 2048 or 4096 keys; smaller tiles give small arrays more
 * workgroups; default 16; the sorted permutation does not depend on it);
 * "fuse_bonds" (1 = each relax pass computes its bonded particles' bond
 * terms in the pass kernel itself, one launch per pass instead of two, for
 * cells whose passes are launch-bound; 0 = a kernel of its own before each
 * pass (default); results do not depend on it). */
sc_status sc_set_option(sc_sim* sim, const char* name, int value);

#ifdef __cplusplus
}
#endif
#endif /* SPATIAL_CELL_H */
