/*
 * md_oracle.h -- CPU FP64 restatement of moldyn's per-step solver path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped.
 *
 * Every function restates the arithmetic of the reference (a Rust crate) in
 * the same operation order, compiled with -ffp-contract=off so that no FMA is
 * formed (Rust never contracts).  Citations are to files of the reference
 * repository AndrewChe7/moldyn.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_lj_potential_and_force, orc_new_lennard_jones
 *       golden digits solver/src/lib.rs:77-89; closed form in long double.
 *   orc_update_force
 *       golden digits solver/src/lib.rs:91-107,109-262; Newton-3 bitwise
 *       antisymmetry; brute-force pair count vs an independent numpy distance
 *       matrix; cell-list variant bitwise equal.
 *   orc_verlet_step (NVE)
 *       golden digits solver/src/lib.rs:109-332, cli/src/tests.rs:35-98;
 *       momentum invariant solver/src/lib.rs:49-75.
 *   orc_verlet_step_cells
 *       bitwise equal to orc_verlet_step over NVE and NPT trajectories of a
 *       liquid, a gas and a two-type system.
 *   macro parameters
 *       golden digits solver/src/lib.rs:264-427; ideal-gas closed form.
 *   Berendsen thermostat / barostat, Nose-Hoover
 *       the reference has no enabled test for them (its only tests are
 *       #[ignore]d, solver/src/lib.rs:429-482); pinned by closed forms
 *       (lambda = 1 at T = T0, mu = 1 at P = P0, one-step hand evaluation
 *       of the published formulas in long double) in the tests.
 *   lattice initialisers
 *       index order solver/src/lib.rs:17-47; box cli/src/tests.rs:9-33.
 *   seeded Maxwell-Boltzmann (no reference counterpart: the reference uses an
 *       unseeded thread_rng) -- pinned by invariants only (exact zero net
 *       momentum by pair negation, sample variance); parity unpinned as to
 *       the draws themselves, which the reference cannot reproduce.
 */
#ifndef MD_ORACLE_H
#define MD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_K_B 1.380648528 /* core/src/lib.rs:15 */

typedef struct {
    double sigma, eps, r_cut, u_cut;
} orc_lj;

/* Type-major particle state, the flattening of State{particles: Vec<Vec<Particle>>}
 * (core/src/particle.rs:5-32).  Arrays are owned by the caller. */
typedef struct {
    int n_types;
    const int64_t* counts; /* particles per type */
    double* pos;           /* n x 3 */
    double* vel;           /* n x 3 */
    double* force;         /* n x 3 */
    double* pot;           /* n     Particle::potential */
    double* temp;          /* n     Particle::temp = sum_j F_ij . r_ij */
    double* mass;          /* n     Particle::mass */
    double box[3];
} orc_state;

typedef struct {
    int kind; /* 0 none, 1 Berendsen, 2 Nose-Hoover */
    double tau, lambda, psi;
} orc_thermostat;

typedef struct {
    int kind; /* 0 none, 1 Berendsen */
    double beta, tau, myu;
} orc_barostat;

/* potential.rs:27-55 */
void orc_new_lennard_jones(double sigma, double eps, orc_lj* out);
/* potential.rs:57-75 */
void orc_lj_potential_and_force(const orc_lj* p, double r, double* u, double* dudr);
/* potential.rs:158-216; table is n_types x n_types, symmetric */
void orc_update_force(orc_state* s, const orc_lj* table, int nthreads);
/* Rows [i0, i1) of orc_update_force for a single-type state (the reference's
 * per-i loop body, potential.rs:176-212): a bounded sample of the O(N^2) work
 * for CPU timing. */
void orc_update_force_rows(orc_state* s, const orc_lj* table, int64_t i0, int64_t i1,
                           int nthreads);
/* Same pair set and per-atom summation order as orc_update_force, candidates
 * found through a periodic cell grid instead of all pairs. */
int orc_update_force_cells(orc_state* s, const orc_lj* table, int nthreads);
/* number of j that contribute to i under potential.rs:178-204 */
void orc_pair_counts(const orc_state* s, const orc_lj* table, double radius_override,
                     int32_t* counts, int nthreads);
/* core/src/particle.rs:120-141 */
void orc_apply_boundary_conditions(orc_state* s);

/* macro_parameters/mod.rs:12-25 */
void orc_center_of_mass_velocity(const orc_state* s, int t, double out[3]);
/* macro_parameters/mod.rs:28-34 */
void orc_momentum(const orc_state* s, int t, double out[3]);
/* macro_parameters/energy.rs:14-22 */
double orc_kinetic_energy(const orc_state* s, int t);
/* macro_parameters/energy.rs:25-37 */
double orc_thermal_energy(const orc_state* s, int t, const double com[3]);
/* macro_parameters/energy.rs:40-49 */
double orc_potential_energy(const orc_state* s, int t);
/* macro_parameters/temperature.rs:4-7 */
double orc_temperature(double thermal_energy, int64_t n);
/* macro_parameters/pressure.rs:5-20 */
double orc_pressure(const orc_state* s, int t, const double com[3]);

/* initializer/thermostat.rs:24-44 */
void orc_thermostat_calculate_lambda(orc_thermostat* th, const orc_state* s, double dt, int t,
                                     double target);
/* initializer/thermostat.rs:47-72 */
void orc_thermostat_update(orc_thermostat* th, orc_state* s, double dt, int t, double target);
/* initializer/barostat.rs:21-36 */
void orc_barostat_calculate_myu(orc_barostat* ba, const orc_state* s, double dt, int t,
                                double target);
/* initializer/barostat.rs:39-56 */
void orc_barostat_update(orc_barostat* ba, orc_state* s, int t);

/* solver/integrator.rs:14-64.  ba/th may be NULL (Option::None). */
void orc_verlet_step(orc_state* s, const orc_lj* table, double dt, orc_barostat* ba,
                     double target_p, orc_thermostat* th, double target_t, int nthreads);
/* The same step with the forces of orc_update_force_cells, for states too
 * large for O(N^2) forces.  Returns 0, or -1 on allocation failure. */
int orc_verlet_step_cells(orc_state* s, const orc_lj* table, double dt, orc_barostat* ba,
                          double target_p, orc_thermostat* th, double target_t, int nthreads);

/* initializer/position.rs:34-51 (UnitCell::U) and :52-99 (UnitCell::FCC) */
void orc_lattice_u(double* pos, const double start[3], const int64_t grid[3], double a);
void orc_lattice_fcc(double* pos, const double start[3], const int64_t grid[3], double a);
/* Seeded Maxwell-Boltzmann with pair negation, structure of
 * initializer/velocity.rs:6-29; draws from SplitMix64 + Box-Muller. */
void orc_mb_velocities(double* vel, int64_t n, double temperature, double mass, uint64_t seed);
/* uniform(0,1) stream used by the jittered-lattice inputs */
void orc_uniform(double* out, int64_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif
