/*
 * sanitize_oracle.c -- the oracle under AddressSanitizer and
 * UndefinedBehaviorSanitizer (TEST INFRASTRUCTURE ONLY, like md_oracle.c).
 *
 * `make -C oracle sanitize SAN_DIR=<dir>` builds md_oracle.c with
 * -fsanitize=address,undefined and this driver into <dir>/oracle_san.  It runs
 * every oracle function on small systems that reach the edge cases of the
 * index arithmetic: one and two particles, a box with fewer than three cells
 * per axis, particles on the box faces, two types, empty types, and
 * NVE / Berendsen NPT / Nose-Hoover trajectories through both force variants,
 * which must agree bitwise.  Any sanitizer finding aborts; a failed check
 * exits 1.  The last line of a clean run is "oracle checks passed".
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "md_oracle.h"

static int g_fail = 0;

#define CHECK(cond, msg)                                                     \
    do {                                                                     \
        if (!(cond)) {                                                       \
            fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, msg);  \
            g_fail = 1;                                                      \
        }                                                                    \
    } while (0)

typedef struct {
    orc_state s;
    int64_t counts[2];
    double *pos, *vel, *force, *pot, *temp, *mass;
} sys_t;

static void sys_alloc(sys_t* y, int n_types, const int64_t* counts, const double* box) {
    int64_t n = 0;
    for (int t = 0; t < n_types; ++t) {
        y->counts[t] = counts[t];
        n += counts[t];
    }
    /* exact sizes, so that any access past n is an ASan finding */
    size_t m = (size_t)(n ? n : 1);
    y->pos = calloc(3 * m, sizeof(double));
    y->vel = calloc(3 * m, sizeof(double));
    y->force = calloc(3 * m, sizeof(double));
    y->pot = calloc(m, sizeof(double));
    y->temp = calloc(m, sizeof(double));
    y->mass = calloc(m, sizeof(double));
    memset(&y->s, 0, sizeof(y->s));
    y->s.n_types = n_types;
    y->s.counts = y->counts;
    y->s.pos = y->pos;
    y->s.vel = y->vel;
    y->s.force = y->force;
    y->s.pot = y->pot;
    y->s.temp = y->temp;
    y->s.mass = y->mass;
    for (int k = 0; k < 3; ++k) y->s.box[k] = box[k];
    int64_t i = 0;
    for (int t = 0; t < n_types; ++t)
        for (int64_t j = 0; j < counts[t]; ++j) y->mass[i++] = t ? 131.29 : 66.335;
}

static void sys_copy(sys_t* dst, const sys_t* src, int64_t n) {
    memcpy(dst->pos, src->pos, 3 * n * sizeof(double));
    memcpy(dst->vel, src->vel, 3 * n * sizeof(double));
    memcpy(dst->force, src->force, 3 * n * sizeof(double));
    memcpy(dst->pot, src->pot, n * sizeof(double));
    memcpy(dst->temp, src->temp, n * sizeof(double));
    memcpy(dst->s.box, src->s.box, sizeof(src->s.box));
}

static void sys_free(sys_t* y) {
    free(y->pos);
    free(y->vel);
    free(y->force);
    free(y->pot);
    free(y->temp);
    free(y->mass);
}

static int same(const double* a, const double* b, int64_t n) { return memcmp(a, b, n * sizeof(double)) == 0; }

/* one system through both force variants and the macro folds */
static void run_case(const char* name, int n_types, const int64_t* counts, const double* box,
                     double jitter, int thermo, int baro, int steps) {
    orc_lj table[4];
    orc_new_lennard_jones(0.3405, 1.654, &table[0]);
    orc_new_lennard_jones(0.3698, 2.015, &table[1]);
    table[2] = table[1];
    orc_new_lennard_jones(0.3963, 2.943, &table[3]);
    sys_t a, b;
    sys_alloc(&a, n_types, counts, box);
    sys_alloc(&b, n_types, counts, box);
    int64_t n = 0;
    for (int t = 0; t < n_types; ++t) n += counts[t];
    double* u = malloc(sizeof(double) * (size_t)(3 * n + 1));
    orc_uniform(u, 3 * n, 11);
    for (int64_t i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            /* a coarse lattice with jitter, some particles exactly on the faces */
            const double side = cbrt((double)n) + 1.0;
            double x = fmod((double)((i * (k + 2)) % (int64_t)side) / side * box[k] + jitter * (u[3 * i + k] - 0.5),
                            box[k]);
            if (x < 0) x += box[k];
            if (i % 7 == 3 && k == 1) x = 0.0;
            a.pos[3 * i + k] = x;
        }
    free(u);
    if (n) orc_mb_velocities(a.vel, n, 150.0, 66.335, 5);
    orc_apply_boundary_conditions(&a.s);
    sys_copy(&b, &a, n);
    orc_update_force(&a.s, table, 2);
    CHECK(orc_update_force_cells(&b.s, table, 2) == 0, "cell forces allocation");
    CHECK(same(a.force, b.force, 3 * n) && same(a.pot, b.pot, n) && same(a.temp, b.temp, n), name);
    int32_t* pc = malloc(sizeof(int32_t) * (size_t)(n ? n : 1));
    orc_pair_counts(&a.s, table, -1.0, pc, 2);
    free(pc);
    if (n_types == 1 && n >= 1) orc_update_force_rows(&a.s, table, 0, n < 2 ? n : 2, 1);
    sys_copy(&b, &a, n);
    orc_thermostat tha = {thermo, 1.0, 0.0, 0.0}, thb = tha;
    orc_barostat baa = {baro ? 1 : 0, 1.0, 50.0, 0.0}, bab = baa;
    for (int s = 0; s < steps; ++s) {
        orc_verlet_step(&a.s, table, 0.002, baro ? &baa : NULL, 1.0, thermo ? &tha : NULL, 120.0, 2);
        CHECK(orc_verlet_step_cells(&b.s, table, 0.002, baro ? &bab : NULL, 1.0, thermo ? &thb : NULL, 120.0, 2) ==
                  0,
              "cell step allocation");
    }
    CHECK(same(a.pos, b.pos, 3 * n) && same(a.vel, b.vel, 3 * n) && same(a.force, b.force, 3 * n), name);
    for (int t = 0; t < n_types; ++t) {
        double com[3], p[3];
        orc_center_of_mass_velocity(&a.s, t, com);
        orc_momentum(&a.s, t, p);
        const double ke = orc_kinetic_energy(&a.s, t), th = orc_thermal_energy(&a.s, t, com);
        const double pe = orc_potential_energy(&a.s, t), pr = orc_pressure(&a.s, t, com);
        const double tk = orc_temperature(th, counts[t]);
        if (counts[t] > 0) CHECK(isfinite(ke) && isfinite(th) && isfinite(pe) && isfinite(pr) && isfinite(tk), name);
        if (thermo == 1 && counts[t] > 0) {
            orc_thermostat_calculate_lambda(&tha, &a.s, 0.002, t, 120.0);
            orc_thermostat_update(&tha, &a.s, 0.002, t, 120.0);
        }
        if (baro && counts[t] > 0) {
            orc_barostat_calculate_myu(&baa, &a.s, 0.002, t, 1.0);
            orc_barostat_update(&baa, &a.s, t);
        }
    }
    sys_free(&a);
    sys_free(&b);
}

int main(void) {
    const double small[3] = {0.9, 0.9, 0.9};   /* 1 cell per axis (r_cut 0.85) */
    const double mid[3] = {2.0, 1.9, 3.1};     /* 2-3 cells per axis */
    const double big[3] = {6.0, 5.0, 4.5};
    const int64_t one[1] = {1}, two[1] = {2}, liquid[1] = {96}, gas[1] = {40};
    const int64_t mixed[2] = {30, 20}, empty_first[2] = {0, 12}, empty_second[2] = {12, 0};
    run_case("one particle", 1, one, small, 0.1, 0, 0, 3);
    run_case("two particles", 1, two, small, 0.1, 1, 1, 3);
    run_case("dense small box", 1, liquid, mid, 0.05, 1, 1, 4);
    run_case("gas big box", 1, gas, big, 0.3, 2, 0, 4);
    run_case("two types", 2, mixed, mid, 0.05, 1, 1, 4);
    run_case("empty first type", 2, empty_first, big, 0.2, 0, 0, 3);
    /* NVE only: the reference's thermostat and barostat divide by a type's
     * particle count, so an empty type gives NaN controls (as in Rust) */
    run_case("empty second type", 2, empty_second, big, 0.2, 0, 0, 3);
    /* lattice and velocity initialisers at exact buffer sizes */
    {
        const double start[3] = {0.0, 0.0, 0.0};
        const int64_t grid[3] = {2, 3, 1};
        double* p = malloc(sizeof(double) * 3 * 6);
        orc_lattice_u(p, start, grid, 0.5);
        free(p);
        p = malloc(sizeof(double) * 3 * 24);
        orc_lattice_fcc(p, start, grid, 0.5);
        free(p);
        for (int64_t n = 0; n < 6; ++n) {
            double* v = malloc(sizeof(double) * (size_t)(3 * n + 1));
            orc_mb_velocities(v, n, 300.0, 39.948, 3);
            free(v);
        }
        double r, dr;
        orc_lj lj;
        orc_new_lennard_jones(0.3405, 1.654, &lj);
        orc_lj_potential_and_force(&lj, 0.1, &r, &dr);
        CHECK(isfinite(r) && isfinite(dr), "close pair");
    }
    if (g_fail) return 1;
    printf("oracle checks passed\n");
    return 0;
}
