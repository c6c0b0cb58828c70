/*
 * cabanapic_amd.h -- C-ABI of the MI355X-native CabanaPIC hot path.
 *
 * One context owns the device state of one simulation (particles, fields,
 * interpolators, accumulators) on the current HIP device and one HIP stream.
 * Entry points are asynchronous on that stream unless they return host data
 * (downloads, energies) or say otherwise.  All functions return CPIC_OK or an
 * error code; cpic_last_error() describes the last failure of the calling
 * thread.  No exception crosses the ABI.
 *
 * Real-valued arguments are passed as double; a float context converts them
 * to float once (exact for values that were computed in float).  Host buffers
 * passed to uploads/downloads hold real_t = float or double according to the
 * context's real_size.
 *
 * Host buffer layouts (identical to the reference's logical layouts):
 *   Ng = (nx+2)(ny+2)(nz+2), voxel = x + (nx+2)*(y + (ny+2)*z)
 *        (src/types.h:184-195); ghost width 1.
 *   fields        9 x Ng, component-major, order ex,ey,ez,cbx,cby,cbz,jfx,
 *                 jfy,jfz (src/types.h:138-170)
 *   interpolators 18 x Ng, component-major, order of src/types.h:64-84
 *   accumulators  Ng x 3 x 4 (jx[4],jy[4],jz[4] per cell,
 *                 src/types.h:109-120)
 *   particles     SoA x,y,z,ux,uy,uz,w (real_t) + int32 cell
 *                 (src/types.h:31-54); x,y,z are in-cell offsets in [-1,1].
 */
#ifndef CABANAPIC_AMD_H
#define CABANAPIC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPIC_OK 0
#define CPIC_ERR_INVALID 1     /* bad argument or input state            */
#define CPIC_ERR_UNSUPPORTED 2 /* e.g. Reflect boundary (fields.h:21-25)  */
#define CPIC_ERR_HIP 3         /* a HIP runtime call failed               */
#define CPIC_ERR_DEVICE 4      /* a kernel flagged an invalid particle    */

/* src/input/deck.h:9-12 */
#define CPIC_BOUNDARY_REFLECT 0
#define CPIC_BOUNDARY_PERIODIC 1

typedef struct cpic_context cpic_context;

typedef struct cpic_config {
  int32_t real_size;     /* 4 = float, 8 = double (REAL_TYPE, types.h:4-8) */
  int32_t num_ghosts;    /* must be 1 (deck.h:267)                         */
  int64_t nx, ny, nz;    /* interior cells per axis                        */
  int64_t max_particles; /* particle capacity of the context               */
  /* Least rows of each cell's particle column (no reference counterpart:
   * the reference stores particles in one flat array); 0 = automatic.  At
   * upload a column holds 2.5x the largest cell count (at least +48 rows,
   * rounded up to a multiple of 64), or every particle of the context where that costs at most 64 MB of
   * columns, and at least column_rows.  Before each push the library
   * checks the previous push's fullest column; within max(8, rows/4) of
   * the capacity it re-lays the columns out on the device with 1.5x the
   * rows (bunching physics: a nonlinear two-stream phase reaches 4x the
   * mean count).  A push that still overflows a column fails with
   * CPIC_ERR_DEVICE. */
  int64_t column_rows;
} cpic_config;

/* Constants of one step, as computed by the driver
 * (example/example.cpp:77-114,179-181). */
typedef struct cpic_step_params {
  double qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp; /* push arguments            */
  double dx, dy, dz, dt;                       /* unload arguments          */
  double px, py, pz, dt_eps0;                  /* field solver arguments    */
  double me, c;                                /* kinetic energy units      */
  int32_t boundary;         /* CPIC_BOUNDARY_*                              */
  int32_t energy_interval;  /* E,B energies every k steps (0 = never);
                               the reference dumps every step
                               (ENERGY_DUMP_INTERVAL, example.cpp:53-55)   */
  int32_t compact_interval; /* particles stay binned by cell; every k-th
                               push also packs each cell's column
                               (0 = every 32nd push)                         */
  int32_t kinetic_interval; /* kinetic energy every k steps (0 = never)     */
} cpic_step_params;

/* ---- context ------------------------------------------------------------ */

/* Replaces the allocations of example/example.cpp:121-144 and the zeroing
 * of Field_Solver's constructor (src/fields.h:279-315) and of
 * initialize_interpolator (src/interpolator.cpp:125-172): every device array
 * starts at zero. */
int cpic_create(const cpic_config *cfg, cpic_context **out);
int cpic_destroy(cpic_context *ctx);
const char *cpic_last_error(void);
int cpic_synchronize(cpic_context *ctx);
/* hipStream_t of the context, for callers that order their own work. */
int cpic_get_stream(cpic_context *ctx, void **stream);

/* ---- data movement (host <-> HBM), synchronous -------------------------- */

/* Particle cells must be interior voxels; anything else is rejected.  The
 * particles are binned into the per-cell columns on the device (the host
 * computes each particle's slot; the arrays pass once over PCIe).  A new
 * column capacity is allocated before the old columns are released, so an
 * allocation failure leaves the resident particles unchanged; a failure
 * after that leaves the context without particles (cpic_num_particles 0).
 * An upload is a fresh particle state: it clears the device error flags. */
int cpic_upload_particles(cpic_context *ctx, int64_t n, const void *x,
                          const void *y, const void *z, const void *ux,
                          const void *uy, const void *uz, const void *w,
                          const int32_t *cell);
int cpic_download_particles(cpic_context *ctx, int64_t n, void *x, void *y,
                            void *z, void *ux, void *uy, void *uz, void *w,
                            int32_t *cell);
/* The upload index of every particle, in the order cpic_download_particles
 * returns them (no reference counterpart: the reference's particles carry no
 * identity; the index lets a caller follow particles through the columns). */
int cpic_download_particle_tags(cpic_context *ctx, int64_t n, int64_t *tags);
int cpic_num_particles(cpic_context *ctx, int64_t *n);
/* Rows of each cell's particle column (no reference counterpart). */
int cpic_column_rows(cpic_context *ctx, int64_t *rows);
/* Grows the particle columns to at least `rows` rows (rounded up to 64) on
 * the device, as the automatic regrowth before a push does when a column
 * nears its capacity: a new column set is allocated and the occupied rows of
 * every column are packed into it in row order.  No-op when the columns
 * already hold that many rows.  No reference counterpart (Cabana's AoSoA is
 * flat); for callers that expect particles to bunch.  Synchronous. */
int cpic_reserve_column_rows(cpic_context *ctx, int64_t rows);
int cpic_upload_fields(cpic_context *ctx, const void *fields);
int cpic_download_fields(cpic_context *ctx, void *fields);
int cpic_upload_interpolators(cpic_context *ctx, const void *interp);
int cpic_download_interpolators(cpic_context *ctx, void *interp);
int cpic_upload_accumulators(cpic_context *ctx, const void *acc);
int cpic_download_accumulators(cpic_context *ctx, void *acc);

/* ---- the hot-path stages (asynchronous) --------------------------------- */

/* load_interpolator_array, src/interpolator.h:14-21,
 * src/interpolator.cpp:4-124 */
int cpic_load_interpolator_array(cpic_context *ctx);

/* clear_accumulator_array, src/accumulator.h:15-21,
 * src/accumulator.cpp:5-43 */
int cpic_clear_accumulator_array(cpic_context *ctx);

/* push<ScatterView> with the inline move_p, src/push.h:7-300,
 * src/move_p.h:59-374: Boris push, cell crossing and charge-conserving
 * deposit, added to the accumulators as the reference's push does.  The
 * first push after an upload and every 32nd one also pack the particle
 * columns (no visible effect but the order within a cell).  boundary must be
 * CPIC_BOUNDARY_PERIODIC. */
int cpic_push(cpic_context *ctx, double qdt_2mc, double cdt_dx,
              double cdt_dy, double cdt_dz, double qsp, int32_t boundary);

/* uncenter_particles, src/uncenter_p.h:4-105, run by example/example.cpp:
 * 204-213 before the step loop when the deck sets perform_uncenter: every
 * particle's momentum takes a backward half rotation in B (qdt_4mc =
 * -0.5 qdt_2mc) and then the forward half E kick u += hax, as the
 * reference writes it (uncenter_p.h:91-93).  Uses the interpolators as
 * loaded; positions and cells are unchanged. */
int cpic_uncenter_particles(cpic_context *ctx, double qdt_2mc);

/* unload_accumulator_array, src/accumulator.h:23-34,
 * src/accumulator.cpp:45-122 */
int cpic_unload_accumulator_array(cpic_context *ctx, double dx, double dy,
                                  double dz, double dt);

/* Field_Solver<EM_Field_Solver>::advance_b, src/fields.h:352-364,668-719
 * (includes the periodic B ghost copy, serial_update_ghosts_B). */
int cpic_advance_b(cpic_context *ctx, double px, double py, double pz);

/* Field_Solver<EM_Field_Solver>::advance_e, src/fields.h:365-378,618-665
 * (includes the periodic J fold, serial_update_ghosts, and J copy). */
int cpic_advance_e(cpic_context *ctx, double px, double py, double pz,
                   double dt_eps0);

/* Field solver used by cpic_advance_b, cpic_advance_e, cpic_energies and
 * cpic_step: the electromagnetic Yee solver (Field_Solver<EM_Field_Solver>,
 * the reference's default) or the electrostatic 1D solver
 * (Field_Solver<ES_Field_Solver_1D>, src/fields.h:467-545, selected in the
 * reference by building with ES_FIELD_SOLVER, example/example.cpp:27-33).
 * ES: advance_b does nothing; advance_e folds the J ghosts as
 * serial_update_ghosts does (ghost cells included) and adds -dt_eps0 J to E
 * on every voxel; the E energy sums every voxel, ghosts included (the B
 * energy is that of the EM solver).  ES needs ny = nz = 1
 * (example/example.cpp:69-74), else CPIC_ERR_UNSUPPORTED.  Default EM. */
#define CPIC_FIELD_SOLVER_EM 0
#define CPIC_FIELD_SOLVER_ES_1D 1
int cpic_set_field_solver(cpic_context *ctx, int32_t solver);

/* Compensated summation of the current deposit (no counterpart in the
 * reference, whose push adds every streak's contributions to the
 * accumulator in real_t, src/push.h:236-254 and src/move_p.h:175-190).
 * enabled = 1: each cell's 12 accumulator sums inside a push pass carry
 * their rounding error (TwoSum) and add it back when the pass writes them;
 * 0 (default): plain real_t sums.  Particles are unaffected; the
 * accumulators differ from the plain push's by rounding only.  Meant for
 * FP32 runs whose field energies are compared with FP64. */
int cpic_set_compensated_deposit(cpic_context *ctx, int32_t enabled);

/* serial_update_ghosts on jfx,jfy,jfz (src/fields.h:102-271) followed by
 * serial_update_ghosts_B on J (src/fields.h:11-100): the ghost update that
 * advance_e performs first, exposed on its own. */
int cpic_update_ghosts_j(cpic_context *ctx);

/* EM_Field_Solver::e_energy / b_energy, src/fields.h:556-615: 0.5 * sum of
 * squares over interior cells with dV = 1, accumulated in FP64.
 * Synchronous. */
int cpic_energies(cpic_context *ctx, double *e_energy, double *b_energy);

/* Kinetic energy sum w me c^2 (gamma-1) in FP64 (the reference has no
 * counterpart).  Synchronous. */
int cpic_kinetic_energy(cpic_context *ctx, double me, double c, double *ke);

/* Gauss's law check in FP64 on the device (the reference's path has none;
 * its charge-conserving deposit, src/move_p.h:156-190, keeps the residual
 * constant under the EM solver).  At every interior node v (the low corner
 * of voxel v):
 *   r(v) = sum_a (e_a(v) - e_a(v - a)) / d_a  -  rho(v) / eps0,
 * with the x edge of voxel v joining nodes v and v + x, the periodic image
 * read for v - a in the low ghost layer, and rho the cloud-in-cell charge
 * density of the particles (weight w qsp, fraction (1 -+ o)/2 per axis to
 * the low/high node, o the in-cell offset), divided by dx dy dz.
 * CPIC_GAUSS_BASELINE stores r as the baseline, out[0] = max |r|;
 * CPIC_GAUSS_CHANGE needs a baseline, out[0] = max |r - baseline|.  Both:
 * out[1] = max |rho / eps0|.  residual (may be NULL) receives r for the
 * nx*ny*nz nodes, x fastest, node (1,1,1) first.  Synchronous. */
#define CPIC_GAUSS_BASELINE 0
#define CPIC_GAUSS_CHANGE 1
int cpic_gauss_residual(cpic_context *ctx, double dx, double dy, double dz,
                        double eps0, double qsp, int32_t mode, double *out,
                        double *residual);

/* Sort of the particles by cell (the commented-out Cabana::sortByKey of
 * example/example.cpp:226-228).  Particles are binned by cell from upload
 * on and every push keeps them binned, so this only validates the context;
 * the order of the particles within one cell is unspecified. */
int cpic_sort_particles(cpic_context *ctx);

/* nsteps iterations of the step loop body, example/example.cpp:216-279:
 * [sort], load_interpolator, clear_accumulator, push, unload,
 * advance_b(1/2), advance_e, advance_b(1/2), energies.  hist (may be NULL)
 * receives one row of 3 doubles (E, B, kinetic) per step; entries of steps
 * without the corresponding dump are left untouched.  The energies stay on
 * the device until the loop ends (one synchronisation per call, where the
 * reference synchronises at every dump).  With particles resident the push
 * computes each cell's interpolator from the fields itself (the same
 * values); the interpolator array is written on the last step, so at return
 * it holds the last push's interpolators.  Synchronous at return. */
int cpic_step(cpic_context *ctx, const cpic_step_params *sp, int64_t nsteps,
              double *hist);

/* cpic_step without the history and without the synchronisation at return:
 * the steps are queued on the context's stream.  Energies the steps take
 * stay on the device, the last ones in the context (cpic_snapshot_begin
 * copies them).  A device error of these steps fails the push two steps
 * later or the next synchronous call.  Stage timing of these steps is
 * collected by the next synchronous call. */
int cpic_step_async(cpic_context *ctx, const cpic_step_params *sp,
                    int64_t nsteps);

/* Per-step dumps without a stall (the partloc / ex1d dumps of
 * example/example.cpp:59-60,130-132,273-277, whose data is that of
 * src/helpers.h:31-70 and src/fields.h:318-350, and the energy line of
 * src/fields.h:741-762).  cpic_snapshot_begin queues, on the context's
 * stream after the work already queued, a gather of: the in-cell offset x,
 * the momentum ux and the voxel of every particle in upload order (the order
 * of the particle arrays the reference dumps); ex at voxels (i, 1, 1),
 * i = 1..nx; and the last E, B and kinetic energies; then a copy of them into
 * pinned host memory the library owns, on a separate copy stream, and
 * returns.  cpic_snapshot_wait blocks until the copy of that slot has landed
 * and fills *out with pointers into it, valid until the next begin on the
 * slot.  Two slots (0, 1) let a caller format one step's dump while the next
 * steps run; a slot must be waited for before it is begun again.  Begin and
 * wait of one slot must not run concurrently; wait may run on another thread
 * than the steps. */
typedef struct cpic_snapshot {
  int64_t n;            /* particles */
  int64_t nx;           /* ex samples */
  int32_t real_size;
  const void *x;        /* in-cell offset x of particle i at [i]          */
  const void *ux;       /* momentum ux                                    */
  const int32_t *cell;  /* voxel                                          */
  const void *ex_line;  /* ex at voxels (1..nx, 1, 1)                     */
  double e_energy, b_energy, kinetic_energy; /* last ones taken (0 if none) */
} cpic_snapshot;
int cpic_snapshot_begin(cpic_context *ctx, int32_t slot);
int cpic_snapshot_wait(cpic_context *ctx, int32_t slot, cpic_snapshot *out);

/* ---- diagnostics --------------------------------------------------------- */

/* Stage timing with HIP events on the context stream.  While enabled, each
 * timed stage launched by cpic_step (and by the single-stage entry points)
 * is bracketed by events; cpic_stage_times returns, per stage, the number
 * of timed launches and their summed milliseconds.  Each event pair delays
 * the stream by a few microseconds.  cpic_set_timing times every stage (or
 * none); cpic_set_timing_stages times the stages whose bit (1 << id) is set.
 * Either call resets the totals.  Stage ids: */
#define CPIC_STAGE_INTERP 0
#define CPIC_STAGE_CLEAR 1
#define CPIC_STAGE_PUSH 2
#define CPIC_STAGE_UNLOAD 3
#define CPIC_STAGE_ADVANCE_B 4
#define CPIC_STAGE_ADVANCE_E 5
#define CPIC_STAGE_ENERGY 6
#define CPIC_STAGE_SORT 7
#define CPIC_NUM_STAGES 8
int cpic_set_timing(cpic_context *ctx, int32_t enabled);
int cpic_set_timing_stages(cpic_context *ctx, uint32_t stages);
int cpic_stage_times(cpic_context *ctx, int64_t *counts, double *ms);

/* Particles that took the move_p branch in the last push, and the device
 * error flags raised since the last particle upload (0 = none).  A push
 * that raises a flag leaves the particle state undefined: later pushes do
 * nothing on the device and cpic_push / cpic_step return CPIC_ERR_DEVICE
 * until particles are uploaded again. */
int cpic_last_push_movers(cpic_context *ctx, int64_t *movers);
int cpic_device_errors(cpic_context *ctx, int32_t *flags);

/* ---- tangent-linear step (forward-mode derivative) ---------------------- */

/* The reference has no derivative of its step; these entry points add one.
 * A tangent context propagates, next to the primal state, its derivative
 * with respect to one deck parameter theta: particles carry d(x, y, z, ux,
 * uy, uz)/dtheta and an FP64 tangent field context carries dE, dB, dJ
 * /dtheta.  theta enters through the initial tangents the caller uploads
 * (for example d(ux)/d(drift) of every particle, or a uniform dB_z = 1 in
 * the tangent fields for the guide field).  The primal step is the one of
 * cpic_step with the push of src/push.h:65-295 and src/move_p.h:93-371 run
 * one thread per particle with atomic deposits: positions, momenta and cells
 * are bit-identical to cpic_push's, accumulators and fields agree up to the
 * order of the floating-point sums.  The derivative is that of the piece of
 * the step each particle is on (face crossings held as taken).
 *
 * cpic_tangent_contexts returns the primal context (fields, interpolators,
 * accumulators in the configured real_t; no particles) and the tangent
 * context (real_size 8): upload/download fields and call cpic_energies on
 * them with the entry points above.  Both use the primal context's stream. */
typedef struct cpic_tangent cpic_tangent;
int cpic_tangent_create(const cpic_config *cfg, cpic_tangent **out);
int cpic_tangent_destroy(cpic_tangent *tg);
int cpic_tangent_contexts(cpic_tangent *tg, cpic_context **primal,
                          cpic_context **tangent);
/* Particles as for cpic_upload_particles (any order, interior cells), and
 * their tangents: 6 x n doubles, tangents of x, y, z, ux, uy, uz in turn. */
int cpic_tangent_upload_particles(cpic_tangent *tg, int64_t n, const void *x,
                                  const void *y, const void *z,
                                  const void *ux, const void *uy,
                                  const void *uz, const void *w,
                                  const int32_t *cell, const double *tangent);
/* Particles in upload order; NULL arrays are skipped. */
int cpic_tangent_download_particles(cpic_tangent *tg, int64_t n, void *x,
                                    void *y, void *z, void *ux, void *uy,
                                    void *uz, void *w, int32_t *cell,
                                    double *tangent);
/* One push with its tangent (adds to both accumulator sets). */
int cpic_tangent_push(cpic_tangent *tg, double qdt_2mc, double cdt_dx,
                      double cdt_dy, double cdt_dz, double qsp,
                      int32_t boundary);
/* nsteps steps of example/example.cpp:216-279 with their tangents.  hist
 * (may be NULL) receives 4 doubles per step: E and B energies
 * (src/fields.h:556-615) and their derivatives sum E.dE, sum B.dB, in
 * FP64.  The energy_interval, compact_interval and kinetic_interval fields
 * of sp are not used.  Synchronous. */
int cpic_tangent_step(cpic_tangent *tg, const cpic_step_params *sp,
                      int64_t nsteps, double *hist);
int cpic_tangent_last_push_movers(cpic_tangent *tg, int64_t *movers);
/* Graph replay of cpic_tangent_step (no reference counterpart): with
 * enabled != 0 (the default) the step is captured once into a HIP graph for
 * its parameters, particle count and history buffer, and replayed once per
 * step, so that a small grid is not bound by kernel launches; 0 launches
 * every stage as it goes.  Steps with stage timing on always launch. */
int cpic_tangent_set_graph(cpic_tangent *tg, int32_t enabled);

/* ---- reverse-mode derivative (adjoint) ---------------------------------- */

/* The reference has no derivative of its step; these entry points add the
 * reverse mode, as a forward()/backward() pair on a primal context.
 * cpic_forward runs nsteps steps of cpic_step on the primal context and
 * keeps its state (particles, fields) every checkpoint_interval steps;
 * cpic_backward then takes the derivative of an objective with respect to
 * the final state (seed_fields: 9 x Ng doubles in the field layout;
 * seed_particles: 6 x n doubles, derivatives with respect to x, y, z, ux,
 * uy, uz of each particle in upload order, or NULL for none) and returns
 * the derivative with respect to the initial state (grad_fields, 9 x Ng;
 * grad_particles, 6 x n in upload order; either may be NULL).  Each segment
 * of checkpoint_interval steps is recomputed from its checkpoint, keeping
 * the state before each of its steps, and transposed step by step: the
 * transposes of src/fields.h:618-719 (Yee updates and ghost handling),
 * src/accumulator.cpp:45-122 (unload), src/interpolator.cpp:4-124 (field
 * gather) and src/push.h:124-256 with src/move_p.h:93-209 (push and
 * deposit, face crossings held as taken), in FP64 whatever real_t is.  It is
 * the transpose of the linearization cpic_tangent_step computes.  A
 * checkpoint re-lays the particle columns out as a later recompute will,
 * so the recompute repeats the forward run bit for bit.  The derivative
 * with respect to a deck parameter is the dot product of the returned
 * initial-state derivative with the initial state's derivative with respect
 * to that parameter.  After cpic_backward the primal context holds the final
 * state again.  Device memory: (ceil(nsteps / K) + 1 + K) snapshots of
 * 32 n + 36 Ng bytes (float; 60 n + 72 Ng in double), see
 * cpic_adjoint_memory.  EM field solver only.  Synchronous. */
typedef struct cpic_adjoint cpic_adjoint;
int cpic_adjoint_create(cpic_context *primal, cpic_adjoint **out);
int cpic_adjoint_destroy(cpic_adjoint *adj);
int cpic_forward(cpic_adjoint *adj, const cpic_step_params *sp, int64_t nsteps,
                 int32_t checkpoint_interval, double *hist);
int cpic_backward(cpic_adjoint *adj, const double *seed_fields,
                  const double *seed_particles, double *grad_fields,
                  double *grad_particles);
/* Bytes of device memory held for the checkpoints (with the final state),
 * for the snapshots of one segment, and for the adjoint arrays. */
int cpic_adjoint_memory(cpic_adjoint *adj, int64_t *checkpoint_bytes,
                        int64_t *segment_bytes, int64_t *adjoint_bytes);

/* ---- allocation-free variant -------------------------------------------- */

/* cpic_energies for caller-owned device fields, with no device allocation:
 * no counterpart in the reference; for callers that cannot call hipMalloc
 * (code under HIP graph capture, serving loops, harnesses that forbid it).
 * fields: device pointer to 9 x Ng values of real_size bytes (the host
 * layout above); out: device pointer to 2 doubles, receiving the E and B
 * energies of src/fields.h:556-615 (0.5 * sum of squares over interior
 * cells, FP64), bit-identical to cpic_energies on the same fields.
 * Asynchronous on stream (a hipStream_t, NULL for the default stream).  The
 * reduction's block partials live in a static 64 KiB array per GPU (at most
 * 4096 blocks, whatever the grid); calls that overlap on one GPU share it
 * and race.  Grids beyond 32-bit voxel indexing are rejected. */
int cpic_field_energies_noalloc(const void *fields, int32_t real_size,
                                int64_t nx, int64_t ny, int64_t nz,
                                double *out, void *stream);

/* ---- deck support (host only, no device call) -------------------------- */

/* Synthetic plasma loader for the decks of the benchmark and test
 * configurations; it plays the role of a deck's Particle_Initializer::init
 * (src/input/deck.h:79-158) for uniform thermal plasmas and beams.
 * Particle p sits in interior cell c = p / nppc (x fastest).  From the
 * counter-based stream r(p,k) = splitmix64(splitmix64(seed) + 8p + k):
 * offsets x,y,z = 2u-1 with u = (r >> 11) 2^-53 (k = 0,1,2); three standard
 * normals n0,n1 (Box-Muller on k = 3,4) and n2 (k = 5,6); momenta
 * u_a = s d_a (1 + pert_amp sin(2 pi pert_mode X / len_x)) + uth_a n_a, where
 * s = +1, or alternately +1/-1 for even/odd p when counter_streaming, and X
 * is the particle's physical x position.  Arithmetic in double, stored as
 * real_t.  n must equal nx*ny*nz*nppc. */
typedef struct cpic_plasma_spec {
  uint64_t seed;
  int64_t nx, ny, nz, nppc;
  double weight;
  double uth[3];
  double drift[3];
  double pert_amp;
  double len_x;
  int32_t pert_mode;
  int32_t counter_streaming;
  /* profile 0: uniform density, drift modulated by
   *   1 + pert_amp sin(2 pi pert_mode x / len_x).
   * profile 1: an electron column along z about the box centre
   *   (len_x/2, len_y/2): weight = weight * g(X, Y) with the ring profile
   *   g = exp(-((r - ring_radius) / ring_width)^2) (1 + pert_amp
   *   cos(pert_mode theta)); the column's field is the periodic solution
   *   of the discrete Poisson equation on the Yee grid for node charge
   *   density column_charge * g minus its mean (the immobile ions
   *   neutralize the box), the same in every z layer; each particle drifts
   *   at E x B / B^2 of that field interpolated at its position as the push
   *   interpolates it, with the guide field (0, 0, guide_bz); u = gamma v
   *   plus the thermal spread. */
  int32_t profile;
  int32_t reserved;
  double len_y;
  double ring_radius, ring_width;
  double column_charge;
  double guide_bz;
} cpic_plasma_spec;
int cpic_init_plasma(const cpic_plasma_spec *spec, int32_t real_size,
                     int64_t n, void *x, void *y, void *z, void *ux, void *uy,
                     void *uz, void *w, int32_t *cell);
/* Initial fields of a profile-1 spec (electron column): ex, ey of the
 * column's periodic field (above) on their Yee edges in every voxel, ghosts
 * included as periodic images (the reference's dioctron deck initializes
 * every voxel, decks/dioctron_3d.cxx:40-64), cbz = guide_bz, all else 0.  fields is
 * [9][Ng] of real_size bytes each, in the reference member order. */
int cpic_init_column_fields(const cpic_plasma_spec *spec, int32_t real_size,
                            void *fields);

#ifdef __cplusplus
}
#endif
#endif
