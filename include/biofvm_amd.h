/*
 * biofvm_amd.h -- C ABI of the MI355X (gfx950) BioFVM diffusion-decay library.
 *
 * One opaque handle (bfa_env) is one microenvironment: a uniform Cartesian
 * grid of nx*ny*nz voxels of edge dx carrying ns substrates, resident on one
 * GPU.  The handle replaces what the reference's LOD_3D_BLOCKING solver reads
 * and writes inside Microenvironment, and the per-agent secretion/uptake loop:
 *
 *   reference entry point                                      replaced by
 *   -----------------------------------------------------------------------------
 *   Microenvironment::simulate_diffusion_decay(dt[,world,cart])  bfa_diffusion_decay
 *     (BioFVM/BioFVM_microenvironment.cpp:1132-1162) dispatching to
 *     diffusion_decay_solver__constant_coefficients_LOD_3D_BLOCKING
 *     (BioFVM/BioFVM_solvers.cpp:102-780) through the function pointers
 *     diffusion_decay_solver / diffusion_decay_solver_mpi
 *     (BioFVM/BioFVM_microenvironment.h:199,206)
 *   Microenvironment::simulate_cell_sources_and_sinks(dt)       bfa_secretion_uptake
 *     (BioFVM/BioFVM_microenvironment.cpp:1370-1384) looping
 *     Basic_Agent::simulate_secretion_and_uptake
 *     (BioFVM/BioFVM_basic_agent.cpp:509-559)
 *   main-loop pair main.cpp:244 then main.cpp:259                bfa_step
 *   Microenvironment::apply_dirichlet_boundaries_conditions       bfa_apply_dirichlet
 *     (BioFVM/BioFVM_microenvironment.cpp:348-446)
 *
 * Every function returns BFA_OK (0) or a nonzero bfa_status and never aborts;
 * bfa_last_error() returns the message of the last failure on that handle.
 * The reference's solver prints and returns on error
 * (BioFVM/BioFVM_solvers.cpp:116-121); the host adapters map a nonzero status
 * to the same print-and-return.  A handle is used by one host thread at a
 * time; all its GPU work runs in order on the handle's own HIP stream.
 */
#ifndef BIOFVM_AMD_H
#define BIOFVM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bfa_env bfa_env;

enum bfa_status {
    BFA_OK = 0,
    BFA_ERR_ARG = 1,    /* invalid argument or size                     */
    BFA_ERR_HIP = 2,    /* a HIP runtime call or kernel launch failed   */
    BFA_ERR_STATE = 3,  /* call made before the state it needs was set  */
    BFA_ERR_NOMEM = 4   /* device allocation failed or budget exceeded  */
};

enum bfa_precision {
    BFA_FP64 = 0,  /* field stored and swept in FP64                          */
    BFA_FP32 = 1   /* field stored and swept in FP32 as scaled deviations
                      u = b + s*v; Thomas precompute and reductions in FP64  */
};

enum bfa_layout {
    BFA_LAYOUT_REF = 0, /* reference order ((i*ny + j)*nz + k)*ns + d, i.e. the
                           p_density_vectors layout of
                           BioFVM/BioFVM_microenvironment.h:169 and
                           BioFVM/BioFVM_mesh.cpp:1157                          */
    BFA_LAYOUT_SOA = 1  /* device order ((d*nz + k)*ny + j)*nx + i            */
};

/* Allocates the device field (zeroed).  Replaces Microenvironment::resize_space
 * (BioFVM/BioFVM_microenvironment.cpp:610-637) for the solver's storage.
 * Edges up to 1024 voxels per axis; ns up to 64. */
int bfa_create(bfa_env **out, int nx, int ny, int nz, int ns, double dx, int precision);
void bfa_destroy(bfa_env *env);
/* Message of the last failure on env; with env == NULL, of the last failed
 * bfa_create in this thread. */
const char *bfa_last_error(const bfa_env *env);

/* Diffusion coefficients D[ns] (length^2/time) and decay rates decay[ns]
 * (1/time): Microenvironment::diffusion_coefficients / decay_rates
 * (BioFVM/BioFVM_microenvironment.h:183-184).  Invalidates the Thomas
 * precompute. */
int bfa_set_coefficients(bfa_env *env, const double *D, const double *decay);

/* Dirichlet state.  active[ns] is dirichlet_activation_vector
 * (BioFVM/BioFVM_microenvironment.cpp:265-273).  face_registered[6] and
 * face_values[6*ns] (face-major, order xmin, xmax, ymin, ymax, zmin, zmax)
 * describe the add_dirichlet_node calls made by initialize_microenvironment
 * (BioFVM/BioFVM_microenvironment.cpp:2649-2786): a boundary voxel takes the
 * values of the last registered face containing it, else 1.0 (:632). */
int bfa_set_dirichlet(bfa_env *env, const uint8_t *active, const uint8_t *face_registered,
                      const double *face_values);
/* Per-voxel Dirichlet values of the boundary voxels: the face-voxel entries of
 * Microenvironment::dirichlet_value_vectors (BioFVM/BioFVM_microenvironment.h,
 * written by add_dirichlet_node / update_dirichlet_node, :218-245) that
 * apply_dirichlet_boundaries_conditions reads per voxel (:348-446).  planes
 * holds six planes back to back, each in reference voxel order with the
 * substrate fastest:
 *   xmin, xmax: [j][k][d]   (ny*nz*ns values each)
 *   ymin, ymax: [i][k][d]   (nx*nz*ns values each)
 *   zmin, zmax: [i][j][d]   (nx*ny*ns values each)
 * A voxel on several faces appears in each of their planes and must carry the
 * same values there (BFA_ERR_ARG otherwise).  The planes replace the face
 * values of bfa_set_dirichlet until called again with planes == NULL; the
 * activation flags stay those of bfa_set_dirichlet. */
int bfa_set_dirichlet_planes(bfa_env *env, const double *planes);

/* Host <-> device field transfer, FP64 on the host in the given layout.  For
 * BFA_FP32 an upload also picks the deviation basis per substrate (b = the
 * Dirichlet value when active, else the field mean; s = a power of two near
 * max|u-b|) unless bfa_set_fp32_basis fixed one. */
int bfa_upload(bfa_env *env, const double *field, int layout);
int bfa_download(bfa_env *env, double *field, int layout);
/* Set every voxel of substrate d to values[d] on the device (no host field
 * transfer), as an upload of that uniform field would (for BFA_FP32 the
 * basis is b = the Dirichlet value when active, else values[d]).  The
 * reference initialises densities per voxel from the substrates' initial
 * conditions (BioFVM/BioFVM_microenvironment.cpp:2487-2520). */
int bfa_fill(bfa_env *env, const double *values);
/* Fix the FP32 basis b[ns], s[ns] (s != 0) before the next upload. */
int bfa_set_fp32_basis(bfa_env *env, const double *b, const double *s);
int bfa_get_fp32_basis(const bfa_env *env, double *b, double *s);

/* Agents in the order the reference's loop visits them (array order).
 * voxel[n] is Basic_Agent::current_voxel_index in the reference numbering
 * (BioFVM/BioFVM_mesh.cpp:1157); agents with active[a] == 0 or voxel < 0 are
 * skipped (BioFVM/BioFVM_basic_agent.cpp:511-512).  c1, c2, e2 are [n][ns]:
 * cell_source_sink_solver_temp1, _temp2 and _temp_export2 as set by
 * Basic_Agent::set_internal_uptake_constants (BioFVM/BioFVM_basic_agent.cpp:
 * 224-274).  n == 0 clears the agents. */
int bfa_set_agents(bfa_env *env, int n, const int32_t *voxel, const uint8_t *active,
                   const double *c1, const double *c2, const double *e2);

/* Which reference solver the step reproduces.  The two differ only in the
 * Dirichlet passes; their recurrences are the same statements at world size 1.
 *   BFA_LOD_3D_BLOCKING: Dirichlet before each sweep and after the z sweep
 *     (diffusion_decay_solver__constant_coefficients_LOD_3D_BLOCKING,
 *     BioFVM/BioFVM_solvers.cpp:345,682,731,778), which
 *     initialize_microenvironment selects (BioFVM_microenvironment.cpp:2506).
 *   BFA_LOD_3D_AVX256D: Dirichlet before each sweep only
 *     (diffusion_decay_solver__constant_coefficients_LOD_3D_AVX256D,
 *     BioFVM/BioFVM_solvers.cpp:1072,1510,1618), the Microenvironment
 *     constructor's default (BioFVM_microenvironment.cpp:143).
 * Default BFA_LOD_3D_BLOCKING. */
enum bfa_solver { BFA_LOD_3D_BLOCKING = 0, BFA_LOD_3D_AVX256D = 1 };
int bfa_set_solver(bfa_env *env, int variant);
int bfa_get_solver(const bfa_env *env);

/* Thomas coefficients across dt changes.  keep_first == 0 (the default)
 * recomputes them whenever dt changes.  keep_first == 1 reproduces the
 * reference: its solver computes them on the first call and keeps them for
 * every later dt (diffusion_solver_setup_done, BioFVM/BioFVM_solvers.cpp:
 * 124-125); coefficient or Dirichlet changes still recompute them, at the dt
 * of that step. */
int bfa_set_dt_caching(bfa_env *env, int keep_first);

/* Internalised-substrate tracking, what Basic_Agent::simulate_secretion_and_uptake
 * does when default_microenvironment_options.
 * track_internalized_substrates_in_each_agent is set
 * (BioFVM/BioFVM_basic_agent.cpp:520-532,553-556): before an agent's update
 *   internalized[a][d] -= W*((1 - c2)*rho + c1)/c2     (W = dx^3, rho the
 *   density the agent sees), and after it internalized[a][d] -= e1[a][d].
 * The totals live on the device, [n][ns] in the caller's agent order of the
 * last bfa_set_agents / bfa_set_agents_state call (n agents), and are updated
 * by every bfa_secretion_uptake and bfa_step.
 * bfa_track_internalized(env, 1, initial, e1) switches tracking on: initial
 * [n][ns] are the current totals (NULL: zeros); e1 [n][ns] is
 * cell_source_sink_solver_temp_export1 (NULL: after bfa_set_agents_state its
 * export_rate * dt, after bfa_set_agents the e1 given before or zeros).
 * enable = 0 switches it off.  A later agent call with the same n keeps the
 * totals (and, from bfa_set_agents, the e1); one with another n switches
 * tracking off.  bfa_get_internalized copies the totals to out [n][ns]. */
int bfa_track_internalized(bfa_env *env, int enable, const double *initial, const double *e1);
int bfa_get_internalized(bfa_env *env, double *out);

/* One simulate_diffusion_decay(dt): Dirichlet, x sweep, Dirichlet, y sweep,
 * Dirichlet, z sweep, Dirichlet (BioFVM/BioFVM_solvers.cpp:345-778); without
 * the last pass under BFA_LOD_3D_AVX256D. */
int bfa_diffusion_decay(bfa_env *env, double dt);
/* One simulate_cell_sources_and_sinks pass with the constants of
 * bfa_set_agents, composed per voxel in ascending agent order. */
int bfa_secretion_uptake(bfa_env *env);
/* nsteps x (bfa_diffusion_decay(dt); bfa_secretion_uptake), replayed from a
 * captured HIP graph. */
int bfa_step(bfa_env *env, double dt, int nsteps);
/* apply_dirichlet_boundaries_conditions on the resident field. */
int bfa_apply_dirichlet(bfa_env *env);

/* Total substrate per species, sum of rho*dx^3 over voxels, FP64 [ns]. */
int bfa_total_mass(bfa_env *env, double *out);

/* Agents from their raw state, with the voxel mapping, the constants and the
 * per-voxel ordering computed on the device: replaces, for the whole agent
 * list, Basic_Agent::update_voxel_index (BioFVM/BioFVM_basic_agent.cpp:177-186)
 * and Basic_Agent::set_internal_uptake_constants (:224-274) followed by
 * bfa_set_agents.  position [n][3]; bbox {xmin, ymin, zmin, xmax, ymax, zmax}
 * of the mesh; volume [n]; secretion, uptake, export_rate (net export) and
 * saturation [n][ns]; active [n] or NULL (all active).  An agent outside the
 * box takes no part, as update_voxel_index deactivates it.  dt is the step the
 * constants are for. */
int bfa_set_agents_state(bfa_env *env, int n, const double *position, const double *bbox,
                         const double *volume, const double *secretion, const double *uptake,
                         const double *export_rate, const double *saturation,
                         const uint8_t *active, double dt);

/* Bulk sources and sinks, Microenvironment::simulate_bulk_sources_and_sinks
 * (BioFVM/BioFVM_microenvironment.cpp:1327-1368).  set_bulk_rates stores the
 * per-value supply rate S, supply target T and uptake rate U (what the
 * reference's bulk_supply_rate_function / bulk_supply_target_densities_function /
 * bulk_uptake_rate_function write, :210-212), [voxel][ns] in BFA_LAYOUT_REF or
 * [ns][N] in BFA_LAYOUT_SOA; all three NULL releases them.  bulk_sources_and_sinks
 * applies rho <- (rho + dt*T*S) / (1 + dt*(U + S)) to the resident field with
 * the reference's roundings; BFA_ERR_STATE if no rates are set. */
int bfa_set_bulk_rates(bfa_env *env, const double *S, const double *T, const double *U, int layout);
int bfa_bulk_sources_and_sinks(bfa_env *env, double dt);

/* Gradient field of every substrate, as Microenvironment::
 * compute_all_gradient_vectors (BioFVM/BioFVM_microenvironment.cpp:1454-1615):
 * central differences (u[+1] - u[-1]) / (2 dx) inside, one-sided differences
 * at the two ends of each axis, 0 along an axis of one node.  Computed on the
 * device from the resident field.  layout BFA_LAYOUT_REF: out[(n*ns + q)*3 + c]
 * with n the reference voxel index (gradient_vectors[n][q][c]);
 * BFA_LAYOUT_SOA: out[(q*3 + c)*N + ((k*ny + j)*nx + i)].  FP64 fields give
 * the reference arithmetic exactly. */
int bfa_gradients(bfa_env *env, double *out, int layout);
/* Gradient at n voxels (reference voxel indices), out[(a*ns + q)*3 + c]:
 * the values compute_all_gradient_vectors gives there, as read through
 * Microenvironment::nearest_gradient_vector (BioFVM_microenvironment.cpp:1443-1452)
 * by chemotaxis.  Only the listed voxels are computed. */
int bfa_gradients_at(bfa_env *env, int n, const int32_t *voxel, double *out);
/* bfa_gradients_at without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same arguments and
 * results as bfa_gradients_at; its device staging (the voxel list and the
 * gradients before the copy back) lives in a static device array of
 * bfa_noalloc_scratch_bytes() bytes per GPU (64 MiB), held until the process
 * ends, instead of per-call allocations: n * (4 + 24 * ns) bytes (the voxel
 * list rounded up to 256 bytes) must fit it, else BFA_ERR_ARG.  Every call on
 * one GPU shares the array, so calls that overlap on one GPU race; this
 * library serialises the calls of one process. */
int bfa_gradients_at_noalloc(bfa_env *env, int n, const int32_t *voxel, double *out);
size_t bfa_noalloc_scratch_bytes(void);
/* Mean device time (ms, HIP events) of the whole-field gradient kernel over
 * reps launches into a device scratch buffer (measurement; no reference
 * counterpart). */
int bfa_time_gradient(bfa_env *env, int reps, double *ms);

/* Wait for all work queued on the handle's stream. */
int bfa_synchronize(bfa_env *env);

/* Measurement helpers.  bfa_time_kernels runs `reps` diffusion+agent steps
 * with one HIP event pair around every kernel on the handle's stream and
 * returns the mean milliseconds per launch in ms[5] = {x sweep, y sweep,
 * z sweep, agents, whole step}.  bfa_copy_kernel times a plain device copy of
 * `bytes` bytes `reps` times (for calibrating HBM byte counters) and returns
 * the mean milliseconds per launch; bfa_copy_kernel_stream does the same with
 * the fastest copy form measured here (one pass, four 16-byte vectors per
 * thread, nontemporal loads and stores; bytes <= 2^40), the achievable-
 * bandwidth reference for the sweeps.  Neither has a counterpart in the
 * reference. */
int bfa_time_kernels(bfa_env *env, double dt, int reps, double *ms);
int bfa_copy_kernel(size_t bytes, int reps, double *ms);
int bfa_copy_kernel_stream(size_t bytes, int reps, double *ms);

/* Kernel shape selection, used by the autotuner: the chunk length L of the
 * in-register line solve per axis (4, 8, 16, 32 or 64, 25 for y and z, 20
 * and 28 for the wave-local x sweep, which takes at most 32; 0 = automatic); the
 * x-sweep form: rows_x == 0 selects the wave-local x sweep (a row's chunks
 * live in one wavefront), rows_x in {1,2,4,...,64} the LDS-tiled x sweep
 * with that many rows per tile; cols_y / cols_z, the x-columns per y / z
 * sweep block (16, 32 or 64; 0 = automatic).  On error the previous shape is
 * kept.  bfa_get_tiles writes {Lx, Ly, Lz, rows_x, cols_y, cols_z}. */
int bfa_set_tiles(bfa_env *env, int Lx, int Ly, int Lz, int rows_x, int cols_y, int cols_z);
int bfa_get_tiles(const bfa_env *env, int *out6);

/* Sweep schedule.  planes == 0: each sweep covers the whole volume.
 * planes > 0: the x and y sweeps run in slabs of `planes` k-planes per
 * substrate, the y sweep of a slab directly after its x sweep (it re-reads
 * the slab from cache), on two streams that overlap neighbouring slabs; the
 * z sweep then covers the volume.  Needs the wave-local x sweep.  In this
 * schedule bfa_time_kernels reports the x+y slab phase in ms[0] and 0 in
 * ms[1]. */
int bfa_set_slab(bfa_env *env, int planes);
int bfa_get_slab(const bfa_env *env);

/* Sweep schedule across substrates: groups == 2 sweeps the first and the
 * second half of the substrates (x, y, z each) on two parallel streams, so
 * that the tail of each kernel overlaps the other half's work; groups == 1
 * (default) sweeps all substrates with one launch per axis.  In bfa_step
 * each half's agent step follows its own z sweep on its stream.  Per-line
 * arithmetic is the same, so the result is bit-identical.  Ignored while a
 * slab schedule is set.  bfa_time_kernels always times the single-group
 * sweeps kernel by kernel; bfa_step uses the schedule.  No reference
 * counterpart (the reference sweeps substrates inside one loop nest,
 * BioFVM/BioFVM_solvers.cpp:352-777). */
int bfa_set_substrate_streams(bfa_env *env, int groups);
int bfa_get_substrate_streams(const bfa_env *env);

/* Grid of the wave-local x sweep: blocks_per_cu 256-thread blocks per compute
 * unit (capped at how many fit), each wavefront striding over row groups; 0
 * (the default) launches as many as fit.  Fewer resident waves can stream
 * faster (at 400^3 x 4 FP32, 4 blocks per CU instead of the 5 that fit);
 * the autotuner measures it.  bfa_get_x_grid reports the requested value
 * and how many fit for the current shape.  No reference counterpart (a
 * launch parameter of the x sweep, BioFVM/BioFVM_solvers.cpp:352-426). */
int bfa_set_x_grid(bfa_env *env, int blocks_per_cu);
int bfa_get_x_grid(const bfa_env *env, int *requested, int *resident);

/* Order in which the y and z sweeps take their column tiles: each of the
 * GPU's 8 XCDs sweeps one contiguous share of the tiles, cut into runs_y
 * (runs_z) sub-runs that it advances together, so that the blocks it holds at
 * one time read from that many separate stretches of the field; 1 (the
 * default) is one contiguous run.  The arithmetic per line is the same, so
 * results are bit-identical for every value.  Range [1, 64]; the autotuner
 * measures it (at 400^3 x 4 FP32 the z sweep is fastest with a few runs).
 * No reference counterpart (a launch parameter of the y and z sweeps,
 * BioFVM/BioFVM_solvers.cpp:427-777). */
int bfa_set_col_runs(bfa_env *env, int runs_y, int runs_z);
int bfa_get_col_runs(const bfa_env *env, int *runs_y, int *runs_z);

/* How the agent step (bfa_step, bfa_secretion_uptake) writes the field:
 * whole_lines = 1 loads and stores the whole 128-byte line around each agent
 * voxel (one lane per line, all the line's agent voxels applied in place),
 * 0 (the default) stores single voxels.  Results are bit-identical.
 * Single-voxel stores leave lines partly written in the caches; their later
 * write-back costs a read-modify-write in memory that the next sweeps pay
 * for (measured: 800^3 x 2 with 10^6 agents, FP64 x sweep 3.51 ms after the
 * agent step against 3.21 ms; 400^3 x 4 FP64 with 10^5 agents, step 2.50
 * against 2.41 ms), while whole lines cost more instructions in the agent
 * step itself (400^3 x 4 FP32 with 10^5 agents: step 1.226 ms with single
 * voxels, 1.238 with whole lines; the small members of a 256-member
 * ensemble solved one by one: 1.11 against 1.02 s).  autotune.autotune_agents
 * times both on the caller's agents.  No reference counterpart (the reference's agent loop
 * updates each voxel in place, BioFVM/BioFVM_basic_agent.cpp:536-552). */
int bfa_set_agent_stores(bfa_env *env, int whole_lines);
int bfa_get_agent_stores(const bfa_env *env);

/* ---- batches of independent microenvironments -------------------------------
 *
 * A batch advances many microenvironments of different shapes and substrate
 * counts together: one persistent kernel per sweep covers every member, and a
 * call of bfa_batch_step replays one captured HIP graph for all its steps.
 * The reference has no counterpart (each Microenvironment is stepped on its
 * own through simulate_diffusion_decay); per member the arithmetic is that of
 * bfa_diffusion_decay + bfa_secretion_uptake.  Members can be added and
 * removed between calls.  One precision per batch.  Member edges: x <= 512,
 * y and z <= 1024.
 */
typedef struct bfa_batch bfa_batch;

int bfa_batch_create(bfa_batch **out, int precision);
void bfa_batch_destroy(bfa_batch *batch);
const char *bfa_batch_last_error(const bfa_batch *batch);
/* Device bytes one member of this shape occupies (for admission control). */
size_t bfa_batch_member_bytes(int nx, int ny, int nz, int ns, int precision);
/* Add a member: grid, dx, coefficients and Dirichlet state as for bfa_create,
 * bfa_set_coefficients and bfa_set_dirichlet.  Writes the member id. */
int bfa_batch_add(bfa_batch *batch, int nx, int ny, int nz, int ns, double dx, const double *D,
                  const double *decay, const uint8_t *active, const uint8_t *face_registered,
                  const double *face_values, int *member_id);
/* Free a member's device memory; its id may be reused by a later add. */
int bfa_batch_remove(bfa_batch *batch, int member_id);
int bfa_batch_upload(bfa_batch *batch, int member_id, const double *field, int layout);
int bfa_batch_download(bfa_batch *batch, int member_id, double *field, int layout);
int bfa_batch_set_agents(bfa_batch *batch, int member_id, int n, const int32_t *voxel,
                         const uint8_t *active, const double *c1, const double *c2,
                         const double *e2);
/* Internalised-substrate tracking of a member's agents, as
 * bfa_track_internalized / bfa_get_internalized for a single microenvironment
 * (BioFVM/BioFVM_basic_agent.cpp:520-532,553-556): totals [n][ns] in the
 * caller's agent order of the member's last bfa_batch_set_agents call
 * (initial NULL: zeros), e1 [n][ns] the net export per step (NULL: kept from
 * the previous call, else zeros).  A later bfa_batch_set_agents call with
 * another agent count switches the member's tracking off.  Members that do
 * not track are unaffected. */
int bfa_batch_track_internalized(bfa_batch *batch, int member_id, int enable, const double *initial,
                                 const double *e1);
int bfa_batch_get_internalized(bfa_batch *batch, int member_id, double *out);
/* Set every voxel of substrate d to values[d] (initial conditions without a
 * host transfer). */
int bfa_batch_fill(bfa_batch *batch, int member_id, const double *values);
/* How many of each following bfa_batch_step call's steps this member takes
 * (it stays untouched for the rest of the call); default: all. */
int bfa_batch_set_member_steps(bfa_batch *batch, int member_id, int steps);
/* nsteps x (diffusion + agents) for every member, asynchronous on the
 * batch's stream. */
int bfa_batch_step(bfa_batch *batch, double dt, int nsteps);
int bfa_batch_synchronize(bfa_batch *batch);
int bfa_batch_total_mass(bfa_batch *batch, int member_id, double *out);
/* Kernel shapes: x chunk length (8, 16, 32), y / z chunk length (4, 8, 16,
 * 32) and columns per tile of the longest lines (16, 32, 64, 128); 0 =
 * automatic.  get writes {Lx, Ly, Lz, cols_y, cols_z}. */
int bfa_batch_set_tiles(bfa_batch *batch, int Lx, int Ly, int Lz, int cols_y, int cols_z);
int bfa_batch_get_tiles(const bfa_batch *batch, int *out5);
/* Threads per block of the y / z column kernels (a multiple of 64, at most
 * 1024, or 512 for chunks of 128 bytes or more); an item whose lines have P
 * chunks then runs as tiles of about threads / P columns.  0 (the default)
 * uses cols x the chunk count of the longest lines.  Rejected (and left
 * unchanged) when the longest lines would get fewer than 16 columns.  get
 * writes the y and z block sizes in use.  No reference counterpart. */
int bfa_batch_set_col_threads(bfa_batch *batch, int threads);
int bfa_batch_get_col_threads(const bfa_batch *batch, int *out2);
/* Stream groups: with G in [2, 4], members (largest first) are dealt
 * round-robin into G groups whose sweeps run on G streams inside the same
 * captured graph, so one group's kernels fill the tails of the others'; 1
 * (the default) runs every member through one kernel per sweep.  Results are
 * bit-identical either way.  No reference counterpart. */
int bfa_batch_set_streams(bfa_batch *batch, int groups);
int bfa_batch_get_streams(const bfa_batch *batch);
/* Fused x + y sweeps (default off): members whose k-plane fits a 128 KiB
 * on-chip buffer are swept along x and y plane by plane in LDS, reading and
 * writing the field once for both sweeps.  Same per-line arithmetic as the
 * separate sweeps except the x chunking (5 nodes).  No reference
 * counterpart: the reference sweeps each axis over the whole grid
 * (BioFVM/BioFVM_solvers.cpp:352-730). */
int bfa_batch_set_fusion(bfa_batch *batch, int enabled);
/* Number of live members that take the fused sweep. */
int bfa_batch_fused_members(const bfa_batch *batch);

#ifdef __cplusplus
}
#endif

#endif /* BIOFVM_AMD_H */
