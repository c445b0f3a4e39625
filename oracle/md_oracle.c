/*
 * md_oracle.c -- CPU FP64 restatement of moldyn's per-step solver path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Build with
 * -O2 -ffp-contract=off (oracle/Makefile).  Operation order follows the Rust
 * reference, citing AndrewChe7/moldyn files; nalgebra's Vector3 arithmetic is
 * component-wise, its dot/norm for 3-vectors is (a*a + b*b) + c*c, and
 * num-traits' pow(6) is powi(6) = s2 * (s2 * s2).
 */
#include "md_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int64_t total_count(const orc_state* s) {
    int64_t n = 0;
    for (int t = 0; t < s->n_types; ++t) n += s->counts[t];
    return n;
}

static int64_t type_offset(const orc_state* s, int t) {
    int64_t o = 0;
    for (int k = 0; k < t; ++k) o += s->counts[k];
    return o;
}

/* potential.rs:57-75: (U - u_cut, dU/dr); zero beyond r_cut. */
void orc_lj_potential_and_force(const orc_lj* p, double r, double* u, double* dudr) {
    if (r > p->r_cut) {
        *u = 0.0;
        *dudr = 0.0;
        return;
    }
    double sr = p->sigma / r;
    double s2 = sr * sr; /* powi(6) by repeated squaring */
    double s4 = s2 * s2;
    double s6 = s2 * s4;
    double s12 = s6 * s6;
    *u = 4.0 * p->eps * (s12 - s6) - p->u_cut;
    *dudr = (24.0 * p->eps / r) * (s6 - 2.0 * s12);
}

/* potential.rs:27-55: r_cut = 2.5 sigma, u_cut = U(r_cut) with u_cut = 0. */
void orc_new_lennard_jones(double sigma, double eps, orc_lj* out) {
    orc_lj p = {sigma, eps, sigma * 2.5, 0.0};
    double u, f;
    orc_lj_potential_and_force(&p, p.r_cut, &u, &f);
    p.u_cut = u;
    *out = p;
}

/* potential.rs:186-200: one conditional minimum-image correction per axis. */
static inline double min_image(double r, double b) {
    if (r < -b / 2.0) {
        r += b;
    } else if (r > b / 2.0) {
        r -= b;
    }
    return r;
}

/* Body of the j loop, potential.rs:181-211, for one (i, j). */
static inline void pair_term(const double* pi, const double* pj, const double* bb, const orc_lj* p,
                             double* f, double* pot, double* temp) {
    double rx = min_image(pj[0] - pi[0], bb[0]);
    double ry = min_image(pj[1] - pi[1], bb[1]);
    double rz = min_image(pj[2] - pi[2], bb[2]);
    double r_abs = sqrt((rx * rx + ry * ry) + rz * rz);
    if (r_abs > p->r_cut) return;
    double u, du;
    orc_lj_potential_and_force(p, r_abs, &u, &du);
    double fx = (rx / r_abs) * du;
    double fy = (ry / r_abs) * du;
    double fz = (rz / r_abs) * du;
    double t = (fx * rx + fy * ry) + fz * rz;
    f[0] += fx;
    f[1] += fy;
    f[2] += fz;
    *pot += u;
    *temp += t;
}

static void zero_forces(orc_state* s) {
    int64_t n = total_count(s);
    memset(s->force, 0, sizeof(double) * 3 * (size_t)n);
    memset(s->pot, 0, sizeof(double) * (size_t)n);
    memset(s->temp, 0, sizeof(double) * (size_t)n);
}

/* potential.rs:158-216.  For each type pair t1 <= t2 only particles of t1
 * accumulate (the reference never adds the t2 <- t1 term); i == j skipped
 * within one type; j ascending. */
void orc_update_force(orc_state* s, const orc_lj* table, int nthreads) {
    zero_forces(s);
    const int T = s->n_types;
    const double* bb = s->box;
    for (int t1 = 0; t1 < T; ++t1) {
        int64_t o1 = type_offset(s, t1), n1 = s->counts[t1];
        for (int t2 = t1; t2 < T; ++t2) {
            int64_t o2 = type_offset(s, t2), n2 = s->counts[t2];
            const orc_lj* p = &table[t1 * T + t2];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
            for (int64_t i = 0; i < n1; ++i) {
                const double* pi = s->pos + 3 * (o1 + i);
                double f[3] = {s->force[3 * (o1 + i)], s->force[3 * (o1 + i) + 1],
                               s->force[3 * (o1 + i) + 2]};
                double pot = s->pot[o1 + i], temp = s->temp[o1 + i];
                for (int64_t j = 0; j < n2; ++j) {
                    if (t1 == t2 && i == j) continue;
                    pair_term(pi, s->pos + 3 * (o2 + j), bb, p, f, &pot, &temp);
                }
                s->force[3 * (o1 + i)] = f[0];
                s->force[3 * (o1 + i) + 1] = f[1];
                s->force[3 * (o1 + i) + 2] = f[2];
                s->pot[o1 + i] = pot;
                s->temp[o1 + i] = temp;
            }
        }
    }
    (void)nthreads;
}

void orc_update_force_rows(orc_state* s, const orc_lj* table, int64_t i0, int64_t i1,
                           int nthreads) {
    const int64_t n = s->counts[0];
    const double* bb = s->box;
    if (i1 > n) i1 = n;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
    for (int64_t i = i0; i < i1; ++i) {
        const double* pi = s->pos + 3 * i;
        double f[3] = {0.0, 0.0, 0.0}, pot = 0.0, temp = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            if (i == j) continue;
            pair_term(pi, s->pos + 3 * j, bb, &table[0], f, &pot, &temp);
        }
        s->force[3 * i] = f[0];
        s->force[3 * i + 1] = f[1];
        s->force[3 * i + 2] = f[2];
        s->pot[i] = pot;
        s->temp[i] = temp;
    }
    (void)nthreads;
}

static int cmp_i64(const void* a, const void* b) {
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

/* Cell-grid variant: candidates from the 27 (deduplicated) neighbouring cells
 * of edge >= max r_cut, evaluated in ascending j so that sums are bitwise
 * those of orc_update_force.  Returns 0, or -1 on allocation failure. */
int orc_update_force_cells(orc_state* s, const orc_lj* table, int nthreads) {
    const int T = s->n_types;
    const int64_t n = total_count(s);
    zero_forces(s);
    if (n == 0) return 0;
    double rmax = 0.0;
    for (int k = 0; k < T * T; ++k)
        if (table[k].r_cut > rmax) rmax = table[k].r_cut;
    int64_t m[3];
    for (int d = 0; d < 3; ++d) {
        double q = floor(s->box[d] / (rmax * (1.0 + 1e-12)));
        m[d] = q < 1.0 ? 1 : (int64_t)q;
        if (m[d] > 1024) m[d] = 1024;
    }
    /* at most about two cells per particle (dilute gases): coarser cells only
     * add candidates, which the criterion rejects, so the result is unchanged */
    while (m[0] * m[1] * m[2] > 2 * n + 27) {
        int d = m[1] > m[0] ? (m[2] > m[1] ? 2 : 1) : (m[2] > m[0] ? 2 : 0);
        m[d] = m[d] - (m[d] / 8 > 1 ? m[d] / 8 : 1);
        if (m[d] < 1) m[d] = 1;
    }
    int64_t ncell = m[0] * m[1] * m[2];
    int64_t* head = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncell + 1));
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* cellof = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int* typeof_ = (int*)malloc(sizeof(int) * (size_t)n);
    if (!head || !order || !cellof || !typeof_) {
        free(head); free(order); free(cellof); free(typeof_);
        return -1;
    }
    {
        int64_t k = 0;
        for (int t = 0; t < T; ++t)
            for (int64_t c = 0; c < s->counts[t]; ++c) typeof_[k++] = t;
    }
    memset(head, 0, sizeof(int64_t) * (size_t)(ncell + 1));
    for (int64_t i = 0; i < n; ++i) {
        int64_t c[3];
        for (int d = 0; d < 3; ++d) {
            double u = s->pos[3 * i + d] / s->box[d];
            u -= floor(u);
            int64_t q = (int64_t)(u * (double)m[d]);
            if (q < 0) q = 0;
            if (q >= m[d]) q = m[d] - 1;
            c[d] = q;
        }
        cellof[i] = (c[0] * m[1] + c[1]) * m[2] + c[2];
        head[cellof[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; ++c) head[c + 1] += head[c];
    {
        int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
        if (!fill) {
            free(head); free(order); free(cellof); free(typeof_);
            return -1;
        }
        memcpy(fill, head, sizeof(int64_t) * (size_t)ncell);
        for (int64_t i = 0; i < n; ++i) order[fill[cellof[i]]++] = i;
        free(fill);
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
#endif
    {
        int64_t cap = 64;
        int64_t* cand = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 256)
#endif
        for (int64_t i = 0; i < n; ++i) {
            int64_t ci = cellof[i];
            int64_t cx = ci / (m[1] * m[2]), cy = (ci / m[2]) % m[1], cz = ci % m[2];
            int64_t nc = 0;
            int64_t offs[3][3];
            int no[3];
            for (int d = 0; d < 3; ++d) {
                int64_t base = d == 0 ? cx : (d == 1 ? cy : cz);
                no[d] = 0;
                for (int64_t o = -1; o <= 1; ++o) {
                    int64_t q = ((base + o) % m[d] + m[d]) % m[d];
                    int dup = 0;
                    for (int k = 0; k < no[d]; ++k) dup |= (offs[d][k] == q);
                    if (!dup) offs[d][no[d]++] = q;
                }
            }
            for (int a = 0; a < no[0]; ++a)
                for (int b = 0; b < no[1]; ++b)
                    for (int c = 0; c < no[2]; ++c) {
                        int64_t cell = (offs[0][a] * m[1] + offs[1][b]) * m[2] + offs[2][c];
                        for (int64_t k = head[cell]; k < head[cell + 1]; ++k) {
                            int64_t j = order[k];
                            if (typeof_[j] < typeof_[i] || j == i) continue;
                            if (nc == cap) {
                                cap *= 2;
                                cand = (int64_t*)realloc(cand, sizeof(int64_t) * (size_t)cap);
                            }
                            cand[nc++] = j;
                        }
                    }
            qsort(cand, (size_t)nc, sizeof(int64_t), cmp_i64);
            double f[3] = {0.0, 0.0, 0.0}, pot = 0.0, temp = 0.0;
            for (int64_t k = 0; k < nc; ++k) {
                int64_t j = cand[k];
                const orc_lj* p = &table[typeof_[i] * T + typeof_[j]];
                pair_term(s->pos + 3 * i, s->pos + 3 * j, s->box, p, f, &pot, &temp);
            }
            s->force[3 * i] = f[0];
            s->force[3 * i + 1] = f[1];
            s->force[3 * i + 2] = f[2];
            s->pot[i] = pot;
            s->temp[i] = temp;
        }
        free(cand);
    }
    free(head); free(order); free(cellof); free(typeof_);
    return 0;
}

/* Count of contributing j per i, criterion of potential.rs:178-204.  With
 * radius_override > 0 that radius replaces every r_cut. */
void orc_pair_counts(const orc_state* s, const orc_lj* table, double radius_override,
                     int32_t* counts, int nthreads) {
    const int T = s->n_types;
    int64_t n = total_count(s);
    memset(counts, 0, sizeof(int32_t) * (size_t)n);
    for (int t1 = 0; t1 < T; ++t1) {
        int64_t o1 = type_offset(s, t1), n1 = s->counts[t1];
        for (int t2 = t1; t2 < T; ++t2) {
            int64_t o2 = type_offset(s, t2), n2 = s->counts[t2];
            double rc = radius_override > 0.0 ? radius_override : table[t1 * T + t2].r_cut;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1)
#endif
            for (int64_t i = 0; i < n1; ++i) {
                const double* pi = s->pos + 3 * (o1 + i);
                int32_t c = 0;
                for (int64_t j = 0; j < n2; ++j) {
                    if (t1 == t2 && i == j) continue;
                    const double* pj = s->pos + 3 * (o2 + j);
                    double rx = min_image(pj[0] - pi[0], s->box[0]);
                    double ry = min_image(pj[1] - pi[1], s->box[1]);
                    double rz = min_image(pj[2] - pi[2], s->box[2]);
                    double r_abs = sqrt((rx * rx + ry * ry) + rz * rz);
                    if (r_abs > rc) continue;
                    ++c;
                }
                counts[o1 + i] += c;
            }
        }
    }
    (void)nthreads;
}

/* core/src/particle.rs:120-141: single conditional wrap per axis. */
void orc_apply_boundary_conditions(orc_state* s) {
    int64_t n = total_count(s);
    for (int64_t i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            double* x = &s->pos[3 * i + d];
            if (*x < 0.0) {
                *x += s->box[d];
            } else if (*x >= s->box[d]) {
                *x -= s->box[d];
            }
        }
    }
}

/* macro_parameters/mod.rs:12-25: fold of (vx, vy, vz, 1) * m, then xyz / w. */
void orc_center_of_mass_velocity(const orc_state* s, int t, double out[3]) {
    int64_t o = type_offset(s, t), n = s->counts[t];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    for (int64_t i = o; i < o + n; ++i) {
        double m = s->mass[i];
        a0 = a0 + s->vel[3 * i] * m;
        a1 = a1 + s->vel[3 * i + 1] * m;
        a2 = a2 + s->vel[3 * i + 2] * m;
        a3 = a3 + 1.0 * m;
    }
    out[0] = a0 / a3;
    out[1] = a1 / a3;
    out[2] = a2 / a3;
}

/* macro_parameters/mod.rs:28-34 */
void orc_momentum(const orc_state* s, int t, double out[3]) {
    int64_t o = type_offset(s, t), n = s->counts[t];
    double p0 = 0.0, p1 = 0.0, p2 = 0.0;
    for (int64_t i = o; i < o + n; ++i) {
        double m = s->mass[i];
        p0 += s->vel[3 * i] * m;
        p1 += s->vel[3 * i + 1] * m;
        p2 += s->vel[3 * i + 2] * m;
    }
    out[0] = p0;
    out[1] = p1;
    out[2] = p2;
}

/* macro_parameters/energy.rs:4-6,14-22: sum of m * (v.v) / 2 */
double orc_kinetic_energy(const orc_state* s, int t) {
    int64_t o = type_offset(s, t), n = s->counts[t];
    double acc = 0.0;
    for (int64_t i = o; i < o + n; ++i) {
        const double* v = s->vel + 3 * i;
        double dot = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
        acc += s->mass[i] * dot / 2.0;
    }
    return acc;
}

/* macro_parameters/energy.rs:8-11,25-37 */
double orc_thermal_energy(const orc_state* s, int t, const double com[3]) {
    int64_t o = type_offset(s, t), n = s->counts[t];
    double acc = 0.0;
    for (int64_t i = o; i < o + n; ++i) {
        const double* v = s->vel + 3 * i;
        double dx = v[0] - com[0], dy = v[1] - com[1], dz = v[2] - com[2];
        double dot = (dx * dx + dy * dy) + dz * dz;
        acc += s->mass[i] * dot / 2.0;
    }
    return acc;
}

/* macro_parameters/energy.rs:40-49 */
double orc_potential_energy(const orc_state* s, int t) {
    int64_t o = type_offset(s, t), n = s->counts[t];
    double acc = 0.0;
    for (int64_t i = o; i < o + n; ++i) acc += s->pot[i];
    return acc / 2.0;
}

/* macro_parameters/temperature.rs:4-7 */
double orc_temperature(double e, int64_t n) {
    double t = (2.0 * e) / (3.0 * (double)n * ORC_K_B);
    return t * 100.0;
}

/* macro_parameters/pressure.rs:5-20 */
double orc_pressure(const orc_state* s, int t, const double com[3]) {
    int64_t o = type_offset(s, t), n = s->counts[t];
    double volume = s->box[0] * s->box[1] * s->box[2];
    double r1 = 0.0, r2 = 0.0;
    for (int64_t i = o; i < o + n; ++i) {
        const double* v = s->vel + 3 * i;
        double m = s->mass[i];
        double dx = v[0] - com[0], dy = v[1] - com[1], dz = v[2] - com[2];
        r1 += m * dx * dx;
        r1 += m * dy * dy;
        r1 += m * dz * dz;
        r2 -= s->temp[i];
    }
    return (r1 + r2 * 0.5) / volume / 3.0;
}

static double type_temperature(const orc_state* s, int t) {
    double com[3];
    orc_center_of_mass_velocity(s, t, com);
    double e = orc_thermal_energy(s, t, com);
    return orc_temperature(e, s->counts[t]);
}

/* initializer/thermostat.rs:24-44 */
void orc_thermostat_calculate_lambda(orc_thermostat* th, const orc_state* s, double dt, int t,
                                     double target) {
    double temperature = type_temperature(s, t);
    if (th->kind == 1) {
        double l2 = 1.0 + dt / th->tau * (target / temperature - 1.0);
        th->lambda = sqrt(l2);
    } else if (th->kind == 2) {
        double psi_dot = -((target / temperature) - 1.0) / th->tau;
        th->psi += psi_dot * (dt / 2.0);
        th->lambda = exp(-th->psi * dt / 2.0);
    }
}

/* initializer/thermostat.rs:47-72 (the temperature is taken before scaling) */
void orc_thermostat_update(orc_thermostat* th, orc_state* s, double dt, int t, double target) {
    double temperature = type_temperature(s, t);
    int64_t o = type_offset(s, t), n = s->counts[t];
    if (th->kind == 1 || th->kind == 2) {
        for (int64_t i = o; i < o + n; ++i) {
            s->vel[3 * i] *= th->lambda;
            s->vel[3 * i + 1] *= th->lambda;
            s->vel[3 * i + 2] *= th->lambda;
        }
    }
    if (th->kind == 2) {
        double psi_dot = -((target / temperature) - 1.0) / th->tau;
        th->psi += psi_dot * (dt / 2.0);
    }
}

/* initializer/barostat.rs:21-36 */
void orc_barostat_calculate_myu(orc_barostat* ba, const orc_state* s, double dt, int t,
                                double target) {
    double com[3];
    orc_center_of_mass_velocity(s, t, com);
    double pressure = orc_pressure(s, t, com);
    if (ba->kind == 1) {
        double m3 = 1.0 + dt * ba->beta / ba->tau * (pressure - target);
        ba->myu = cbrt(m3);
    }
}

/* initializer/barostat.rs:39-56 */
void orc_barostat_update(orc_barostat* ba, orc_state* s, int t) {
    if (ba->kind != 1) return;
    s->box[0] *= ba->myu;
    s->box[1] *= ba->myu;
    s->box[2] *= ba->myu;
    int64_t o = type_offset(s, t), n = s->counts[t];
    for (int64_t i = o; i < o + n; ++i) {
        s->pos[3 * i] *= ba->myu;
        s->pos[3 * i + 1] *= ba->myu;
        s->pos[3 * i + 2] *= ba->myu;
    }
}

/* solver/integrator.rs:14-64; cells selects orc_update_force_cells for the
 * force evaluation (bitwise the same forces).  Returns 0, or -1 when the cell
 * grid could not be allocated (the state is then mid-step). */
static int verlet_step_impl(orc_state* s, const orc_lj* table, double dt, orc_barostat* ba,
                            double target_p, orc_thermostat* th, double target_t, int nthreads,
                            int cells) {
    const int T = s->n_types;
    if (ba)
        for (int t = 0; t < T; ++t) orc_barostat_calculate_myu(ba, s, dt, t, target_p);
    if (th)
        for (int t = 0; t < T; ++t) orc_thermostat_calculate_lambda(th, s, dt, t, target_t);
    for (int t = 0; t < T; ++t) {
        int64_t o = type_offset(s, t), n = s->counts[t];
        if (n == 0) continue;
        double c = dt / (2.0 * s->mass[o]);
        for (int64_t i = o; i < o + n; ++i) {
            s->vel[3 * i] = s->vel[3 * i] + s->force[3 * i] * c;
            s->vel[3 * i + 1] = s->vel[3 * i + 1] + s->force[3 * i + 1] * c;
            s->vel[3 * i + 2] = s->vel[3 * i + 2] + s->force[3 * i + 2] * c;
        }
    }
    if (th)
        for (int t = 0; t < T; ++t) orc_thermostat_update(th, s, dt, t, target_t);
    {
        int64_t n = total_count(s);
        for (int64_t i = 0; i < 3 * n; ++i) s->pos[i] += s->vel[i] * dt;
    }
    orc_apply_boundary_conditions(s);
    if (cells) {
        if (orc_update_force_cells(s, table, nthreads) != 0) return -1;
    } else {
        orc_update_force(s, table, nthreads);
    }
    for (int t = 0; t < T; ++t) {
        int64_t o = type_offset(s, t), n = s->counts[t];
        if (n == 0) continue;
        double c = dt / (2.0 * s->mass[o]);
        for (int64_t i = o; i < o + n; ++i) {
            s->vel[3 * i] += s->force[3 * i] * c;
            s->vel[3 * i + 1] += s->force[3 * i + 1] * c;
            s->vel[3 * i + 2] += s->force[3 * i + 2] * c;
        }
    }
    if (ba)
        for (int t = 0; t < T; ++t) orc_barostat_update(ba, s, t);
    return 0;
}

void orc_verlet_step(orc_state* s, const orc_lj* table, double dt, orc_barostat* ba,
                     double target_p, orc_thermostat* th, double target_t, int nthreads) {
    (void)verlet_step_impl(s, table, dt, ba, target_p, th, target_t, nthreads, 0);
}

int orc_verlet_step_cells(orc_state* s, const orc_lj* table, double dt, orc_barostat* ba,
                          double target_p, orc_thermostat* th, double target_t, int nthreads) {
    return verlet_step_impl(s, table, dt, ba, target_p, th, target_t, nthreads, 1);
}

/* initializer/position.rs:34-51: index x*ny*nz + y*nz + z. */
void orc_lattice_u(double* pos, const double start[3], const int64_t grid[3], double a) {
    for (int64_t x = 0; x < grid[0]; ++x)
        for (int64_t y = 0; y < grid[1]; ++y)
            for (int64_t z = 0; z < grid[2]; ++z) {
                int64_t k = x * grid[1] * grid[2] + y * grid[2] + z;
                pos[3 * k] = start[0] + (double)x * a;
                pos[3 * k + 1] = start[1] + (double)y * a;
                pos[3 * k + 2] = start[2] + (double)z * a;
            }
}

/* initializer/position.rs:52-99: four sites per cell, index 4*cell + k. */
void orc_lattice_fcc(double* pos, const double start[3], const int64_t grid[3], double a) {
    for (int64_t x = 0; x < grid[0]; ++x)
        for (int64_t y = 0; y < grid[1]; ++y)
            for (int64_t z = 0; z < grid[2]; ++z) {
                int64_t c = x * grid[1] * grid[2] + y * grid[2] + z;
                double X = (double)x, Y = (double)y, Z = (double)z;
                double* p = pos + 12 * c;
                p[0] = start[0] + X * a;
                p[1] = start[1] + Y * a;
                p[2] = start[2] + Z * a;
                p[3] = start[0] + X * a;
                p[4] = start[1] + (Y + 0.5) * a;
                p[5] = start[2] + (Z + 0.5) * a;
                p[6] = start[0] + (X + 0.5) * a;
                p[7] = start[1] + Y * a;
                p[8] = start[2] + (Z + 0.5) * a;
                p[9] = start[0] + (X + 0.5) * a;
                p[10] = start[1] + (Y + 0.5) * a;
                p[11] = start[2] + Z * a;
            }
}

/* SplitMix64 (Steele, Lea, Flood 2014). */
static uint64_t splitmix64(uint64_t* st) {
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double unit_open(uint64_t* st) {
    return ((double)(splitmix64(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_uniform(double* out, int64_t n, uint64_t seed) {
    uint64_t st = seed;
    for (int64_t i = 0; i < n; ++i) out[i] = unit_open(&st);
}

/* Box-Muller: each pair of uniforms yields r cos(2 pi u2) then r sin(2 pi u2). */
typedef struct {
    uint64_t st;
    int have;
    double spare;
} normal_gen;

static double next_normal(normal_gen* g) {
    if (g->have) {
        g->have = 0;
        return g->spare;
    }
    double u1 = unit_open(&g->st), u2 = unit_open(&g->st);
    double r = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    g->spare = r * sin(th);
    g->have = 1;
    return r * cos(th);
}

/* initializer/velocity.rs:6-29 structure: sigma = sqrt(K_B * (T * 0.01) / m),
 * particle i and i + n/2 get opposite velocities; odd n leaves the last at 0. */
void orc_mb_velocities(double* vel, int64_t n, double temperature, double mass, uint64_t seed) {
    double tp = temperature * 0.01;
    double sigma = sqrt(ORC_K_B * tp / mass);
    normal_gen g = {seed, 0, 0.0};
    int64_t h = n / 2;
    memset(vel, 0, sizeof(double) * 3 * (size_t)n);
    for (int64_t i = 0; i < h; ++i) {
        double x = sigma * next_normal(&g);
        double y = sigma * next_normal(&g);
        double z = sigma * next_normal(&g);
        vel[3 * i] = x;
        vel[3 * i + 1] = y;
        vel[3 * i + 2] = z;
        vel[3 * (i + h)] = -x;
        vel[3 * (i + h) + 1] = -y;
        vel[3 * (i + h) + 2] = -z;
    }
}
