/*
 * moldyn_hip.h -- C ABI of the MI355X (gfx950) molecular-dynamics step engine.
 *
 * Drop-in for the per-step solver path of moldyn (AndrewChe7/moldyn): the
 * integrator, Lennard-Jones potential, Berendsen / Nose-Hoover thermostat,
 * Berendsen barostat and macro-parameter functions of the `moldyn_solver`
 * crate.  The reference is a Rust crate without an FFI layer; every entry
 * point below names the Rust item whose behaviour it reproduces, so that a
 * Rust `extern "C"` block (or ctypes, or cgo) can bind it unchanged.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Positions/velocities/forces are N x 3
 *    double arrays (x, y, z per particle), particles ordered type-major as in
 *    State{particles: Vec<Vec<Particle>>} (core/src/particle.rs:26-32);
 *    for a batch of replicas, replica-major, then type-major.
 *  - Units are the reference's: nm, ps, 1e-27 kg, zJ, MPa, K (K_B below).
 *  - Every call returns 0 on success or a negative MD_ERR_* code; the text of
 *    the last error is md_last_error().  Nothing aborts the process (the
 *    reference panics on the same conditions).
 *  - A context owns its device memory and one HIP stream; one host thread
 *    drives one context at a time.  Contexts are independent.
 */
#ifndef MOLDYN_HIP_H
#define MOLDYN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MD_K_B 1.380648528 /* core/src/lib.rs:15 */

enum {
    MD_OK = 0,
    MD_ERR_ARG = -1,          /* bad argument (null pointer, size, id) */
    MD_ERR_HIP = -2,          /* HIP runtime failure */
    MD_ERR_NOMEM = -3,        /* device allocation failed */
    MD_ERR_UNIMPLEMENTED = -4,/* Custom variants: the reference's todo!() */
    MD_ERR_BAROSTAT = -5,     /* mu^3 <= 0 or non-finite T / P */
    MD_ERR_OVERFLOW = -6,     /* neighbour list or cell capacity could not be grown */
    MD_ERR_NO_DEVICE = -7     /* no usable gfx950 device */
};

/* Force evaluation paths (md_options.force_mode). */
enum {
    MD_FORCE_FP64_ALLPAIRS = 0, /* O(N^2), reference criterion and sum order: bitwise update_force */
    MD_FORCE_FP64_LIST = 1,     /* cell grid + Verlet list, FP64: same pair set and sum order */
    MD_FORCE_FP32_LIST = 2,     /* cell grid + Verlet list, FP32 pair forces, FP64 reductions */
    MD_FORCE_MLP = 3            /* learned per-atom energy surrogate (one particle type), see md_mlp_set_params */
};

/* Thermostat, initializer/thermostat.rs:4-20 (enum Thermostat) plus the
 * target temperature the reference passes beside it (integrator.rs:15).
 * kind: 0 none, 1 Berendsen{tau, lambda}, 2 NoseHoover{tau, psi, lambda},
 * 3 Custom (rejected with MD_ERR_UNIMPLEMENTED, as the reference's todo!()).
 * lambda and psi are in/out, mutated like the Rust enum fields. */
typedef struct {
    int kind;
    double tau;
    double target_t; /* K */
    double lambda;
    double psi;
} md_thermostat;

/* Barostat, initializer/barostat.rs:4-16 (enum Barostat) plus target pressure.
 * kind: 0 none, 1 Berendsen{beta, tau, myu}, 2 Custom (MD_ERR_UNIMPLEMENTED).
 * myu is in/out. */
typedef struct {
    int kind;
    double beta;
    double tau;
    double target_p; /* program units (MPa) */
    double myu;
} md_barostat;

/* Engine options (no reference counterpart: the reference has one O(N^2) path). */
typedef struct {
    int force_mode;        /* MD_FORCE_*; default MD_FORCE_FP32_LIST */
    double skin;           /* Verlet skin, nm; default 0.3 * sigma(0,0) */
    int block_size;        /* threads per block of the per-atom kernels: 64, 128, 256 or 512; 0 = automatic
                              (256, or 512 for MD_FORCE_MLP); default 256 */
    int list_capacity;     /* initial neighbours per atom; 0 = automatic; grown on overflow */
    int cell_capacity;     /* initial atoms per cell slot array; 0 = automatic; grown on overflow */
    double cells_per_atom; /* cell-grid budget per replica, cells per atom; default 2 */
    int graph_chunk;       /* launches per captured hipGraph (rounded up to even); 0 = default 32, 1 = no graphs */
    int tiles_per_block;   /* particle tiles (block_size slots each) per block of the per-particle
                              kernels; 0 = automatic (about four blocks per CU on the list paths) */
} md_options;

typedef struct md_ctx md_ctx;

/* Lennard-Jones scalar, potential.rs:57-75 Potential::get_potential_and_force:
 * returns U(r) - u_cut and dU/dr, both 0 when r > r_cut.  Host-side. */
void md_lj_potential_and_force(double sigma, double eps, double r_cut, double u_cut, double r,
                               double* u, double* dudr);
/* potential.rs:27-55 Potential::new_lennard_jones: r_cut = 2.5 sigma, u_cut = U(r_cut). */
void md_new_lennard_jones(double sigma, double eps, double* r_cut, double* u_cut);

/* Context for one system: initializer/position.rs:116-137 initialize_particles
 * (counts per type) + the masses the ParticleDatabase holds
 * (core/src/particles_database.rs:67-79) + State::boundary_box. */
int md_create(md_ctx** out, int n_types, const int64_t* counts, const double* masses,
              const double box[3]);
/* Batch of independent replicas in one context: counts is n_replicas x n_types,
 * boxes n_replicas x 3; masses per type shared.  No reference counterpart
 * (the reference runs one State at a time). */
int md_create_replicas(md_ctx** out, int n_replicas, int n_types, const int64_t* counts,
                       const double* masses, const double* boxes);
void md_destroy(md_ctx* ctx);
const char* md_last_error(const md_ctx* ctx);
/* Text for an error raised before a context exists (md_create failures). */
const char* md_global_error(void);

int md_get_options(const md_ctx* ctx, md_options* out);
int md_set_options(md_ctx* ctx, const md_options* opt);

/* PotentialsDatabase::set_potential(t0, t1, Potential::new_lennard_jones(sigma, eps))
 * (potential.rs:141-144, 27-55).  Unset pairs use the default Ar LJ(0.3418, 1.712)
 * (potential.rs:95-101). */
int md_set_pair_potential(md_ctx* ctx, uint16_t t0, uint16_t t1, double sigma, double eps);
/* Same with all four serialized fields, as read from potentials.json
 * (potential.rs:104-139: Potential::LennardJones{sigma, eps, r_cut, u_cut}). */
int md_set_pair_potential_lj(md_ctx* ctx, uint16_t t0, uint16_t t1, double sigma, double eps,
                             double r_cut, double u_cut);

/* Replace positions, velocities and box (save_data.rs:129-151 StateToSave -> State):
 * forces, potentials and virials are zeroed like a freshly loaded State.
 * box is 3 doubles per replica. */
int md_upload_state(md_ctx* ctx, const double* x, const double* v, const double* box);
/* Hand over the force-side fields of a State (Particle::force, ::potential,
 * ::temp, core/src/particle.rs:11-16) after md_upload_state, so that a step
 * continues from a state whose forces the caller already holds
 * (integrator.rs:28-34 reads them in the first half kick).  Any pointer may
 * be NULL (that field stays zero). */
int md_upload_forces(md_ctx* ctx, const double* force, const double* pot, const double* virial);
/* Copy back; any output pointer may be NULL.  F is N x 3, pot and virial are
 * Particle::potential and Particle::temp of the last force evaluation. */
int md_download_state(md_ctx* ctx, double* x, double* v, double* force, double* pot,
                      double* virial, double* box);

/* update_force(&PotentialsDatabase, &mut State), potential.rs:158-216. */
int md_update_force(md_ctx* ctx);
/* State::apply_boundary_conditions, core/src/particle.rs:120-141. */
int md_apply_boundary_conditions(md_ctx* ctx);

/* n_steps x Integrator::VerletMethod.calculate(db, state, dt, barostat, thermostat)
 * (solver/integrator.rs:14-64).  th / ba may be NULL (Option::None); their
 * lambda, psi and myu fields are updated in place as the Rust enums are. */
int md_step(md_ctx* ctx, double dt, md_thermostat* th, md_barostat* ba, int64_t n_steps);
/* Same for a replica batch: th and ba are arrays of n_replicas (or NULL). */
int md_step_replicas(md_ctx* ctx, double dt, md_thermostat* th, md_barostat* ba,
                     int64_t n_steps);

/* Thermostat::calculate_lambda / update (initializer/thermostat.rs:24-72) and
 * Barostat::calculate_myu / update (initializer/barostat.rs:21-56) on the
 * device-resident state of replica 0, for callers that drive the integrator
 * themselves.  dt, type_id and target have the Rust meaning. */
int md_thermostat_calculate_lambda(md_ctx* ctx, md_thermostat* th, double dt, uint16_t type_id);
int md_thermostat_update(md_ctx* ctx, md_thermostat* th, double dt, uint16_t type_id);
int md_barostat_calculate_myu(md_ctx* ctx, md_barostat* ba, double dt, uint16_t type_id);
int md_barostat_update(md_ctx* ctx, md_barostat* ba, uint16_t type_id);

/* Macro parameters of (replica, type) for the current state (macro_parameters/):
 * out[0] kinetic energy   (energy.rs:14-22)
 * out[1] potential energy (energy.rs:40-49, sum / 2)
 * out[2] thermal energy   (energy.rs:25-37)
 * out[3] temperature K    (temperature.rs:4-7, per-type count)
 * out[4] pressure         (pressure.rs:5-20)
 * out[5..7] momentum      (mod.rs:28-34)
 * out[8..10] centre-of-mass velocity (mod.rs:12-25) */
int md_macro(md_ctx* ctx, int replica, uint16_t type_id, double out[11]);

/* No counterpart in the reference: md_macro for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  The state
 * of one particle type is the caller's device memory: v (n x 3), pot and
 * virial (n, Particle::potential / Particle::temp; either may be NULL for
 * zeros); mass is the type's mass and box[3] (host) the boundary box.  The
 * eleven values of md_macro are written to `out`, 11 doubles of device
 * memory, by kernels launched on `stream` (a hipStream_t; NULL the null
 * stream); the call neither synchronises nor copies to the host, and
 * allocates nothing: its partial sums live in a fixed 106 KB array of
 * device data per GPU, held until the process ends.  n must lie in
 * [1, 2^40] (MD_ERR_ARG otherwise, message in md_global_error).  Calls
 * that overlap on one GPU share that array and race. */
int md_macro_noalloc(const double* v, const double* pot, const double* virial, int64_t n, double mass,
                     const double box[3], double* out, void* stream);

/* Lattice initialisers (host-side), initializer/position.rs:34-99:
 * kind 0 UnitCell::U (grid[0]*grid[1]*grid[2] sites), 1 UnitCell::FCC (4 per cell). */
int md_init_lattice(int kind, const double start[3], const int64_t grid[3], double a, double* pos);
/* Seeded Maxwell-Boltzmann velocities with pair negation, structure of
 * initializer/velocity.rs:6-29 (the reference's thread_rng is unseeded; this
 * uses SplitMix64 + Box-Muller so inputs are reproducible). */
int md_init_mb_velocities(double* vel, int64_t n, double temperature, double mass, uint64_t seed);

/* The same initialisers on the device, into particle group (replica,
 * type_id) of a context (a context that never held a state first gets
 * initialize_particles' blank one: positions and velocities 0).
 * md_init_positions: UnitCell::initialize_particles_position for
 * start + (x, y, z) a over grid[0] x grid[1] x grid[2] cells
 * (position.rs:24-104), with the checks of position.rs:163-171 and :35, :53
 * (MD_ERR_ARG "OutOfBoundary" when grid[k] a exceeds the box, "TooBig" when
 * the sites outnumber the group); particles past the sites keep their place.
 * md_init_velocities: initialize_velocities_maxwell_boltzmann
 * (velocity.rs:6-29) with the draws of md_init_mb_velocities, one thread per
 * particle pair (equal to it up to the last bits of log, sin and cos).
 * Either resets the controls as md_upload_state does; call md_update_force
 * before stepping, as cli/src/commands.rs:102 does. */
int md_init_positions(md_ctx* ctx, int replica, uint16_t type_id, int kind, const double start[3],
                      const int64_t grid[3], double a);
int md_init_velocities(md_ctx* ctx, int replica, uint16_t type_id, double temperature, uint64_t seed);

/* Learned per-atom energy surrogate (no reference counterpart: the reference
 * has Lennard-Jones only).  With force_mode MD_FORCE_MLP, md_update_force and
 * md_step use E_i = e_scale * MLP(x_i) + e_shift with
 *   x_i = (G_i - g_mean) * g_inv_std,
 *   G_ik = sum_{j, r_ij < 2.5 sigma} exp(-(r_ij - mu_k)^2 / (2 w^2)) (cos(pi r_ij / 2.5 sigma) + 1) / 2,
 *   mu_k = sigma (0.8 + 1.7 k / 31), w = 0.1 sigma, k = 0..31 (sigma of type pair (0, 0)),
 *   MLP = 32-128-128-1 with SiLU, BF16 weights and activations, FP32 accumulation,
 * and forces F_i = -dE/dr_i.  theta holds MD_MLP_NUM_PARAMS floats:
 *   W1 [128][32] | b1 [128] | W2 [128][128] | b2 [128] | w3 [128] | b3.
 * g_lo / g_hi (nullable): per-feature range seen in training; a launch in
 * which any feature leaves it uses the LJ forces instead (counted as a
 * fallback).  judge_every > 0: every that many steps the LJ forces of the
 * same positions are computed beside and the force RMSE recorded. */
#define MD_MLP_NUM_PARAMS 20865
int md_mlp_set_params(md_ctx* ctx, const float* theta, const float* g_mean, const float* g_inv_std,
                      const float* g_lo, const float* g_hi, double e_scale, double e_shift,
                      int judge_every);
/* out[0] relative force RMSE (vs LJ) at the last check, out[1] largest over
 * the checks, out[2] checks, out[3] fallback steps, out[4] RMS LJ force
 * component at the last check. */
int md_mlp_get_stats(md_ctx* ctx, int replica, double out[5]);

/* Surrogate training on the device (no reference counterpart).  A data set
 * of n_frames periodic frames of n_atoms atoms each: positions (frame-major,
 * N x 3), boxes (3 per frame), per-atom energy labels e_ref (the LJ
 * Particle::potential / 2) and forces f_ref.  The features of every frame are
 * computed once at creation (all pairs, basis of md_mlp_set_params with the
 * given sigma).  Training runs FP32 kernels: forward, analytic force
 * backward, loss L = w_e mean (E - e_ref)^2 + w_f mean (F - f_ref)^2 and its
 * gradient through the forces (double backward), Adam. */
typedef struct md_trainer md_trainer;
int md_trainer_create(md_trainer** out, int n_frames, int n_atoms, const double* pos, const double* boxes,
                      const double* e_ref, const double* f_ref, double sigma);
void md_trainer_destroy(md_trainer* tr);
const char* md_trainer_last_error(const md_trainer* tr);
/* Per-feature mean, standard deviation, minimum and maximum over the data set
 * (out: 4 x 32). */
int md_trainer_feature_stats(md_trainer* tr, double out[128]);
/* Parameters (layout of md_mlp_set_params) and normalisation; resets Adam. */
int md_trainer_set_params(md_trainer* tr, const float* theta, const float* g_mean, const float* g_inv_std,
                          double e_scale, double e_shift);
int md_trainer_get_params(md_trainer* tr, float* theta);
/* Loss over the whole data set: out = {L, energy MSE, force MSE per component};
 * grad (nullable) receives dL/dtheta.  batch_frames bounds the work arrays. */
int md_trainer_loss(md_trainer* tr, double w_e, double w_f, int batch_frames, double out[3], float* grad);
/* One epoch: minibatches of batch_frames frames (in the given order, or
 * sequential when order is NULL), one Adam step (beta 0.9, 0.999) per
 * minibatch on its mean loss; out = mean over the minibatches. */
int md_trainer_epoch(md_trainer* tr, int batch_frames, double lr, double w_e, double w_f, const int32_t* order,
                     double out[3]);
/* Per-atom energies and forces of one frame under the current parameters. */
int md_trainer_predict(md_trainer* tr, int frame, double* e_out, double* f_out);

/* Observability (no reference counterpart). */
typedef struct {
    int64_t steps;          /* real steps taken in the last md_step call */
    int64_t launches;       /* step-kernel launches (includes list-refresh-only launches) */
    int64_t rebuilds;       /* neighbour-list rebuilds in the last call */
    int64_t refresh_steps;  /* launches that only refilled the cell grid */
    int64_t retries;        /* capacity-overflow replays */
    int32_t list_capacity;
    int32_t cell_capacity;
    double mean_list_len;   /* at the last rebuild */
    int32_t max_list_len;
    int32_t n_slots;        /* padded particle slots on the device */
    int64_t grows;          /* capacity increases between chunks (no replay) */
    int64_t list_visits;    /* neighbour-list entries read by the step kernel while profiling */
    int64_t split_steps;    /* steps whose drift exceeded the skin (drift, rebuild, then forces) */
} md_stats;
int md_get_stats(md_ctx* ctx, md_stats* out);
/* Per-kernel device time of the launches of md_step: the step kernel
 * (kernel 0; the pair-force kernel on the surrogate path), the finalize /
 * list-build kernel (1), the all-pairs force kernel (2), the drift kernel
 * (3; the first half step on the surrogate path), the surrogate's network
 * kernel (5) and its feature kernel (6); kernel 4 reports nothing.
 * Every block records its start and end on the GPU wall clock
 * (s_memrealtime), so the captured graphs keep running unchanged.
 *   total_ms / launches: stream time, from the end of the previous kernel of
 *     the launch sequence to the end of this one, over launches that ran back
 *     to back (the interval a dispatch trace such as rocprofv3 reports);
 *   span_ms / span_launches: first block start to last block end.
 * Any output may be NULL.  md_set_profiling(ctx, 1) also clears the totals. */
int md_set_profiling(md_ctx* ctx, int on);
int md_kernel_time(md_ctx* ctx, int kernel, double* total_ms, int64_t* launches, double* span_ms,
                   int64_t* span_launches);

/* ---------------------------------------------------------------- state files
 * The reference's on-disk state (core/src/save_data.rs): per state n,
 * data/<n>.csv with the header
 *   id,position_x,position_y,position_z,velocity_x,velocity_y,velocity_z
 * and one row per particle (types in order, id = type), and row n of bb.csv
 * ("x,y,z" header) holding the box.  Every f64 is written as the reference's
 * csv 1.3 crate writes it (the ryu crate's shortest round-trip text), so the
 * files are byte-for-byte the reference's.  Files are created or truncated
 * (the reference opens without truncation, which leaves stale bytes behind
 * a shorter rewrite; on new files the output is the same). */

/* Asynchronous writer of one replica of a context: md_frame_writer_save(w, n)
 * replaces StateToSave::from(&state).save_to_file(path, n)
 * (save_data.rs:192-213, called at cli/src/commands.rs:183-184,189-190).
 * The text is made on the GPU on the context's stream (so it holds the state
 * of that moment); copies and file writes run on a thread of the writer while
 * the caller goes on, at most two states in flight (save waits for a free
 * slot).  md_frame_writer_flush waits for every write and returns the first
 * error.  The context must outlive the writer.  stats: frames, text bytes,
 * GPU ms, copy ms, file-write ms, ms that save waited for a slot. */
typedef struct md_frame_writer md_frame_writer;
int md_frame_writer_create(md_frame_writer** out, md_ctx* ctx, const char* dir, int replica);
int md_frame_writer_save(md_frame_writer* w, int64_t state_number);
int md_frame_writer_flush(md_frame_writer* w);
int md_frame_writer_stats(md_frame_writer* w, double out[6]);
const char* md_frame_writer_last_error(md_frame_writer* w);
void md_frame_writer_destroy(md_frame_writer* w);

/* StateToSave::load_from_file(path, n) (save_data.rs:215-229): rows of
 * data/<n>.csv in file order (columns found by header name) and row n of
 * bb.csv into box (nullable).  With ids, x and v NULL only *n_rows is set;
 * otherwise cap must hold *n_rows.  Errors: md_global_error(). */
int md_frame_load(const char* dir, int64_t state_number, int64_t cap, int64_t* n_rows, uint16_t* ids, double* x,
                  double* v, double* box);

/* GPU formatter of numeric CSV tables in the same text: rows of an optional
 * u64 key column and ncol (1..10) f64 columns, "," between fields and "\n"
 * after each row.  md_text_table writes the rows (no header) to out (NULL:
 * only *len is set); md_csv_write_table writes header, "\n" and the rows to
 * path, or an empty file for no rows, as the csv crate does.  Used for
 * bb.csv, macro.csv (DataFileMacro::save_to_file, save_data.rs:291-300) and
 * the numbers of db.csv and potentials.json. */
typedef struct md_text md_text;
int md_text_create(md_text** out);
void md_text_destroy(md_text* h);
const char* md_text_last_error(const md_text* h);
int md_text_table(md_text* h, int64_t n_rows, int ncol, const uint64_t* keys, const double* vals, char* out,
                  int64_t cap, int64_t* len);
int md_csv_write_table(md_text* h, const char* path, const char* header, int64_t n_rows, int ncol,
                       const uint64_t* keys, const double* vals);

#ifdef __cplusplus
}
#endif
#endif
