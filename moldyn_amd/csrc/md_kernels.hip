// md_kernels.hip -- gfx950 kernels of the MD step engine.
//
// Per time step of a replica the engine runs two launches:
//   k_step      one thread per particle slot: half kick, thermostat scale,
//               drift, wrap (integrator.rs:28-45), Verlet-list force
//               (potential.rs:158-216) with the neighbours' drift recomputed
//               from the input buffer, second half kick and barostat scale
//               (integrator.rs:47-58), per-block FP64 partial sums.  When the
//               list needs refreshing the same launch rebuilds it at entry
//               from the cell grid filled by the previous launch, and fills
//               the other grid buffer at exit.
//   k_finalize  one block per replica: fixed-order reduction of the block
//               partials, then the Berendsen / Nose-Hoover lambda, psi and
//               Berendsen mu of the next step (thermostat.rs:24-72,
//               barostat.rs:21-56), box rescale and the list-validity
//               decision.  Everything stays on the device; the host only
//               replays a captured graph.
// The O(N^2) FP64 path (k_drift + k_allpairs) restates update_force exactly:
// same pair criterion, same operation order, j ascending.
//
// All code is built with -ffp-contract=off: the FP64 paths must not form FMAs
// the reference (Rust) never forms.  The FP32 pair kernel uses explicit fmaf.
#include <math.h>

#include "md_device.h"

namespace md {

__device__ __forceinline__ double min_image(double r, double b) {
    // potential.rs:186-200: one correction against b / 2
    if (r < -b / 2.0) {
        r += b;
    } else if (r > b / 2.0) {
        r -= b;
    }
    return r;
}

__device__ __forceinline__ double wrap1(double x, double b) {
    // core/src/particle.rs:124-138
    if (x < 0.0) {
        x += b;
    } else if (x >= b) {
        x -= b;
    }
    return x;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block reduction of NS columns in a fixed order (wave butterfly, then waves
// in index order); the columns set in mask are stored to out[k * stride].
template <int BS>
__device__ __forceinline__ void block_store(const double (&v)[NS], double* out, int64_t stride,
                                            uint32_t mask) {
    constexpr int NW = BS / 64;
    __shared__ double red[NW * NS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        if (!((mask >> k) & 1u)) continue;
        const double t = k >= kFirstMax ? wave_max(v[k]) : wave_sum(v[k]);
        if (lane == 0) red[w * NS + k] = t;
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < NS && ((mask >> k) & 1u)) {
        double acc = red[k];
        for (int i = 1; i < NW; ++i)
            acc = k >= kFirstMax ? fmax(acc, red[i * NS + k]) : acc + red[i * NS + k];
        out[k * stride] = acc;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- pair terms

// potential.rs:201-211 with LJ of potential.rs:57-69, FP64, reference order.
__device__ __forceinline__ void pair_exact(double dx, double dy, double dz, double r2,
                                           const PairParam& p, double& fx, double& fy,
                                           double& fz, double& pot, double& vir) {
    const double r = sqrt(r2);
    const double sr = p.sigma / r;
    const double s2 = sr * sr;
    const double s4 = s2 * s2;
    const double s6 = s2 * s4;
    const double s12 = s6 * s6;
    const double u = 4.0 * p.eps * (s12 - s6) - p.u_cut;
    const double du = (24.0 * p.eps / r) * (s6 - 2.0 * s12);
    const double gx = (dx / r) * du;
    const double gy = (dy / r) * du;
    const double gz = (dz / r) * du;
    const double t = (gx * dx + gy * dy) + gz * dz;
    fx += gx;
    fy += gy;
    fz += gz;
    pot += u;
    vir += t;
}

// Same physics in FP32 from the FP64 separation: F/r = 24 eps (s6 - 2 s12) / r^2,
// virial F.r = (F/r) r^2, U = 4 eps (s12 - s6) - u_cut.
__device__ __forceinline__ void pair_fp32(double dx, double dy, double dz, const PairParam& p,
                                          float& fx, float& fy, float& fz, float& pot,
                                          float& vir) {
    const float ax = (float)dx, ay = (float)dy, az = (float)dz;
    const float r2 = fmaf(ax, ax, fmaf(ay, ay, az * az));
    const float ir2 = 1.0f / r2;
    const float s2 = p.sig2f * ir2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    const float fr = p.eps24f * ir2 * fmaf(-2.0f, s12, s6);
    fx = fmaf(ax, fr, fx);
    fy = fmaf(ay, fr, fy);
    fz = fmaf(az, fr, fz);
    pot += fmaf(p.eps4f, s12 - s6, -p.ucutf);
    vir = fmaf(fr, r2, vir);
}

template <typename FT>
struct Acc;
template <>
struct Acc<double> {
    double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0, vir = 0.0;
    __device__ __forceinline__ void add(double dx, double dy, double dz, double r2,
                                        const PairParam& p) {
        pair_exact(dx, dy, dz, r2, p, fx, fy, fz, pot, vir);
    }
};
template <>
struct Acc<float> {
    float fx = 0.0f, fy = 0.0f, fz = 0.0f, pot = 0.0f, vir = 0.0f;
    __device__ __forceinline__ void add(double dx, double dy, double dz, double,
                                        const PairParam& p) {
        pair_fp32(dx, dy, dz, p, fx, fy, fz, pot, vir);
    }
};

// ---------------------------------------------------------------- cell grid

__device__ __forceinline__ int cell_coord(double x, double b, int m) {
    double u = x / b;
    u = u - floor(u);
    int q = (int)(u * (double)m);
    return q < 0 ? 0 : (q >= m ? m - 1 : q);
}

__device__ __forceinline__ void grid_insert(const DevState& d, const RepCtrl& C,
                                            const RepInfo& R, int64_t slot, double x, double y,
                                            double z) {
    const uint32_t ep = C.ins_epoch;
    const int buf = ep & 1;
    const int* gd = C.gdim[buf];
    const double* gb = C.gbox[buf];
    const int64_t cell = R.cell_base + ((int64_t)cell_coord(x, gb[0], gd[0]) * gd[1] +
                                        cell_coord(y, gb[1], gd[1])) * gd[2] +
                         cell_coord(z, gb[2], gd[2]);
    unsigned long long* h = d.cell_hdr[buf] + cell;
    unsigned long long old = __hip_atomic_load(h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long assumed;
    uint32_t k;
    do {
        assumed = old;
        k = ((uint32_t)(assumed >> 32) == ep) ? (uint32_t)assumed : 0u;
        const unsigned long long nw = ((unsigned long long)ep << 32) | (unsigned long long)(k + 1u);
        old = atomicCAS(h, assumed, nw);
    } while (old != assumed);
    if ((int)k < d.cell_cap) {
        d.cell_slots[buf][cell * d.cell_cap + k] = (int32_t)slot;
    } else {
        atomicOr(&d.stats->cell_overflow, 1u);
        atomicMax(&d.stats->max_cell, (int)k + 1);
    }
}

// Verlet list of `slot` from grid `fill_epoch`: every j of the same replica,
// type >= own type (potential.rs:171-180), j != slot, within rlist under the
// reference minimum image, sorted ascending so that sums follow the
// reference's j order.
template <bool MULTI>
__device__ void build_list(const DevState& d, const RepCtrl& C, const RepParams& P,
                           const RepInfo& R, int64_t slot, int ti, const double* xs,
                           const double* ys, const double* zs, double x, double y, double z) {
    const uint32_t ep = C.fill_epoch;
    const int buf = ep & 1;
    const int* gd = C.gdim[buf];
    const double* gb = C.gbox[buf];
    const int cx = cell_coord(x, gb[0], gd[0]);
    const int cy = cell_coord(y, gb[1], gd[1]);
    const int cz = cell_coord(z, gb[2], gd[2]);
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double rl2 = P.rlist * P.rlist;
    const int cap = d.list_cap;
    const int cc = d.cell_cap;
    const int64_t N = d.n_slots;
    int32_t* __restrict__ col = d.nbr + slot;
    const unsigned long long* __restrict__ hdr = d.cell_hdr[buf];
    const int32_t* __restrict__ cs = d.cell_slots[buf];
    int n = 0;
    const int lx = gd[0] >= 3 ? -1 : 0, hx = gd[0] >= 2 ? 1 : 0;
    const int ly = gd[1] >= 3 ? -1 : 0, hy = gd[1] >= 2 ? 1 : 0;
    const int lz = gd[2] >= 3 ? -1 : 0, hz = gd[2] >= 2 ? 1 : 0;
    for (int ox = lx; ox <= hx; ++ox) {
        const int qx = (cx + ox + gd[0]) % gd[0];
        for (int oy = ly; oy <= hy; ++oy) {
            const int qy = (cy + oy + gd[1]) % gd[1];
            for (int oz = lz; oz <= hz; ++oz) {
                const int qz = (cz + oz + gd[2]) % gd[2];
                const int64_t cell = R.cell_base + ((int64_t)qx * gd[1] + qy) * gd[2] + qz;
                const unsigned long long h = hdr[cell];
                if ((uint32_t)(h >> 32) != ep) continue;
                int cnt = (int)(uint32_t)h;
                cnt = cnt < cc ? cnt : cc;
                for (int k = 0; k < cnt; ++k) {
                    const int32_t j = cs[cell * cc + k];
                    if ((int64_t)j == slot) continue;
                    if (MULTI && (int)d.slot_type[j] < ti) continue;
                    const double dx = min_image(xs[j] - x, bx);
                    const double dy = min_image(ys[j] - y, by);
                    const double dz = min_image(zs[j] - z, bz);
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (!(r2 <= rl2)) continue;
                    if (n < cap) {
                        int p = n;
                        while (p > 0 && col[(int64_t)(p - 1) * N] > j) {
                            col[(int64_t)p * N] = col[(int64_t)(p - 1) * N];
                            --p;
                        }
                        col[(int64_t)p * N] = j;
                    }
                    ++n;
                }
            }
        }
    }
    d.ncount[slot] = n < cap ? n : cap;
    if (n > cap) {
        atomicOr(&d.stats->list_overflow, 1u);
        atomicMax(&d.stats->max_list, n);
    }
}

// ---------------------------------------------------------------- fused step

template <typename FT, int BS, bool MULTI>
__global__ __launch_bounds__(BS) void k_step(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const int r = G.rep;
    const RepCtrl& C = d.ctrl[r];
    const uint32_t flags = C.flags;
    if (flags & F_EXIT) return;
    const RepParams& P = d.params[r];
    const RepInfo R = d.reps[r];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    const bool valid = (s - G.start) < G.count;
    const int cur = C.cur, nx = cur ^ 1;
    double* const part_out = d.partial + b;
    const int64_t pstride = d.n_blocks;
    if (flags & F_NULL) {
        if (valid) grid_insert(d, C, R, s, d.x[cur][s], d.y[cur][s], d.z[cur][s]);
        return;
    }
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double dt = P.dt, lam = C.lambda, mu = C.myu;
    const double m = G.mass;
    const int ti = G.type;
    const double c = P.cm[ti];
    const uint32_t kick_cols = P.th_kind == 2 ? kColsKick : 0u;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (flags & F_SPLIT_A) {
        // first half of a split step: kick, scale, drift, wrap into cur ^ 1 and
        // fill the grid with the drifted positions (the drift exceeds the skin)
        if (valid) {
            const double vkx = d.vx[cur][s] + (double)((const FT*)d.fx[cur])[s] * c;
            const double vky = d.vy[cur][s] + (double)((const FT*)d.fy[cur])[s] * c;
            const double vkz = d.vz[cur][s] + (double)((const FT*)d.fz[cur])[s] * c;
            part[S_KVX] = vkx * m;
            part[S_KVY] = vky * m;
            part[S_KVZ] = vkz * m;
            part[S_KV2] = m * ((vkx * vkx + vky * vky) + vkz * vkz);
            const double v1x = vkx * lam, v1y = vky * lam, v1z = vkz * lam;
            const double x1 = wrap1(d.x[cur][s] + v1x * dt, bx);
            const double y1 = wrap1(d.y[cur][s] + v1y * dt, by);
            const double z1 = wrap1(d.z[cur][s] + v1z * dt, bz);
            d.x[nx][s] = x1;
            d.y[nx][s] = y1;
            d.z[nx][s] = z1;
            d.vx[nx][s] = v1x;
            d.vy[nx][s] = v1y;
            d.vz[nx][s] = v1z;
            part[S_VMAX] = (v1x * v1x + v1y * v1y) + v1z * v1z;
            grid_insert(d, C, R, s, x1, y1, z1);
        }
        block_store<BS>(part, part_out, pstride, kick_cols | kColsDrift);
        return;
    }
    // second half of a split step: drifted state in cur ^ 1, results to cur
    const bool split = (flags & F_SPLIT_B) != 0;
    const int src = split ? nx : cur;
    const int dst = split ? cur : nx;
    const double* xs = d.x[src];
    const double* ys = d.y[src];
    const double* zs = d.z[src];
    const double* vxs = d.vx[src];
    const double* vys = d.vy[src];
    const double* vzs = d.vz[src];
    const FT* fxs = (const FT*)d.fx[cur];
    const FT* fys = (const FT*)d.fy[cur];
    const FT* fzs = (const FT*)d.fz[cur];
    if (valid) {
        double x1, y1, z1, v1x, v1y, v1z;
        if (split) {
            x1 = xs[s];
            y1 = ys[s];
            z1 = zs[s];
            v1x = vxs[s];
            v1y = vys[s];
            v1z = vzs[s];
            build_list<MULTI>(d, C, P, R, s, ti, xs, ys, zs, x1, y1, z1);
        } else {
            const double x = xs[s], y = ys[s], z = zs[s];
            if (flags & F_REBUILD) build_list<MULTI>(d, C, P, R, s, ti, xs, ys, zs, x, y, z);
            // integrator.rs:28-34 first half kick
            const double vkx = vxs[s] + (double)fxs[s] * c;
            const double vky = vys[s] + (double)fys[s] * c;
            const double vkz = vzs[s] + (double)fzs[s] * c;
            part[S_KVX] = vkx * m;
            part[S_KVY] = vky * m;
            part[S_KVZ] = vkz * m;
            part[S_KV2] = m * ((vkx * vkx + vky * vky) + vkz * vkz);
            // thermostat.rs:55-57 scale, integrator.rs:40-45 drift + wrap
            v1x = vkx * lam;
            v1y = vky * lam;
            v1z = vkz * lam;
            x1 = wrap1(x + v1x * dt, bx);
            y1 = wrap1(y + v1y * dt, by);
            z1 = wrap1(z + v1z * dt, bz);
            part[S_VMAX] = (v1x * v1x + v1y * v1y) + v1z * v1z;
        }
        // potential.rs:158-216 over the Verlet list at the drifted positions;
        // a neighbour's drift is recomputed from the input buffer with the
        // same operations its own thread performs, so it is bitwise the same
        Acc<FT> acc;
        const int nn = d.ncount[s];
        const PairParam p0 = d.pairs[0];
        for (int k = 0; k < nn; ++k) {
            const int32_t j = d.nbr[(int64_t)k * d.n_slots + s];
            int tj = 0;
            double cj = c;
            if (MULTI) {
                tj = d.slot_type[j];
                cj = P.cm[tj];
            }
            double xj, yj, zj;
            if (split) {
                xj = xs[j];
                yj = ys[j];
                zj = zs[j];
            } else {
                xj = wrap1(xs[j] + ((vxs[j] + (double)fxs[j] * cj) * lam) * dt, bx);
                yj = wrap1(ys[j] + ((vys[j] + (double)fys[j] * cj) * lam) * dt, by);
                zj = wrap1(zs[j] + ((vzs[j] + (double)fzs[j] * cj) * lam) * dt, bz);
            }
            const double dx = min_image(xj - x1, bx);
            const double dy = min_image(yj - y1, by);
            const double dz = min_image(zj - z1, bz);
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            const PairParam& pp = MULTI ? d.pairs[ti * d.n_types + tj] : p0;
            if (r2 <= pp.r2max) acc.add(dx, dy, dz, r2, pp);
        }
        // integrator.rs:47-53 second half kick, barostat.rs:45-48 scale
        const double v2x = v1x + (double)acc.fx * c;
        const double v2y = v1y + (double)acc.fy * c;
        const double v2z = v1z + (double)acc.fz * c;
        const double x2 = x1 * mu, y2 = y1 * mu, z2 = z1 * mu;
        d.x[dst][s] = x2;
        d.y[dst][s] = y2;
        d.z[dst][s] = z2;
        d.vx[dst][s] = v2x;
        d.vy[dst][s] = v2y;
        d.vz[dst][s] = v2z;
        ((FT*)d.fx[dst])[s] = acc.fx;
        ((FT*)d.fy[dst])[s] = acc.fy;
        ((FT*)d.fz[dst])[s] = acc.fz;
        if (flags & F_WRITE_PV) {
            d.pot[s] = (double)acc.pot;
            d.vir[s] = (double)acc.vir;
        }
        part[S_MVX] = v2x * m;
        part[S_MVY] = v2y * m;
        part[S_MVZ] = v2z * m;
        part[S_MV2] = m * ((v2x * v2x + v2y * v2y) + v2z * v2z);
        part[S_POT] = (double)acc.pot;
        part[S_VIR] = (double)acc.vir;
        part[S_WV] = (v2x * v2x + v2y * v2y) + v2z * v2z;
        part[S_WA] = (((double)acc.fx * (double)acc.fx + (double)acc.fy * (double)acc.fy) +
                      (double)acc.fz * (double)acc.fz) * P.inv_m2[ti];
        if (flags & F_INSERT) grid_insert(d, C, R, s, x2, y2, z2);
    }
    // a split step stored its kick / drift columns in the first half
    block_store<BS>(part, part_out, pstride, split ? kColsState : (kColsState | kick_cols | kColsDrift));
}

// Force evaluation at the current positions through a freshly built list
// (update_force for the list paths).  Grid `fill_epoch` must hold x[cur].
template <typename FT, int BS, bool MULTI>
__global__ __launch_bounds__(BS) void k_force_eval(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const int r = G.rep;
    const RepCtrl& C = d.ctrl[r];
    const RepParams& P = d.params[r];
    const RepInfo R = d.reps[r];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    const bool valid = (s - G.start) < G.count;
    const int cur = C.cur;
    const double* xs = d.x[cur];
    const double* ys = d.y[cur];
    const double* zs = d.z[cur];
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double m = G.mass;
    const int ti = G.type;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (valid) {
        const double x = xs[s], y = ys[s], z = zs[s];
        build_list<MULTI>(d, C, P, R, s, ti, xs, ys, zs, x, y, z);
        Acc<FT> acc;
        const int nn = d.ncount[s];
        for (int k = 0; k < nn; ++k) {
            const int32_t j = d.nbr[(int64_t)k * d.n_slots + s];
            const int tj = MULTI ? (int)d.slot_type[j] : 0;
            const double dx = min_image(xs[j] - x, bx);
            const double dy = min_image(ys[j] - y, by);
            const double dz = min_image(zs[j] - z, bz);
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            const PairParam& pp = d.pairs[ti * d.n_types + tj];
            if (r2 <= pp.r2max) acc.add(dx, dy, dz, r2, pp);
        }
        ((FT*)d.fx[cur])[s] = acc.fx;
        ((FT*)d.fy[cur])[s] = acc.fy;
        ((FT*)d.fz[cur])[s] = acc.fz;
        d.pot[s] = (double)acc.pot;
        d.vir[s] = (double)acc.vir;
        const double vx = d.vx[cur][s], vy = d.vy[cur][s], vz = d.vz[cur][s];
        part[S_MVX] = vx * m;
        part[S_MVY] = vy * m;
        part[S_MVZ] = vz * m;
        part[S_MV2] = m * ((vx * vx + vy * vy) + vz * vz);
        part[S_POT] = (double)acc.pot;
        part[S_VIR] = (double)acc.vir;
        part[S_WV] = (vx * vx + vy * vy) + vz * vz;
        part[S_WA] = (((double)acc.fx * (double)acc.fx + (double)acc.fy * (double)acc.fy) +
                      (double)acc.fz * (double)acc.fz) / (m * m);
    }
    block_store<BS>(part, d.partial + b, d.n_blocks, kColsAll);
}

template <int BS>
__global__ __launch_bounds__(BS) void k_insert_all(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = d.ctrl[G.rep];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    if ((s - G.start) >= G.count) return;
    const int cur = C.cur;
    grid_insert(d, C, d.reps[G.rep], s, d.x[cur][s], d.y[cur][s], d.z[cur][s]);
}

// Partial sums of the current state (after an upload or a host-side change).
template <typename FT, int BS>
__global__ __launch_bounds__(BS) void k_reduce_state(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = d.ctrl[G.rep];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    const bool valid = (s - G.start) < G.count;
    const int cur = C.cur;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (valid) {
        const double m = G.mass;
        const double vx = d.vx[cur][s], vy = d.vy[cur][s], vz = d.vz[cur][s];
        part[S_MVX] = vx * m;
        part[S_MVY] = vy * m;
        part[S_MVZ] = vz * m;
        part[S_MV2] = m * ((vx * vx + vy * vy) + vz * vz);
        part[S_POT] = d.pot[s];
        part[S_VIR] = d.vir[s];
        const double fx = (double)((const FT*)d.fx[cur])[s];
        const double fy = (double)((const FT*)d.fy[cur])[s];
        const double fz = (double)((const FT*)d.fz[cur])[s];
        part[S_WV] = (vx * vx + vy * vy) + vz * vz;
        part[S_WA] = ((fx * fx + fy * fy) + fz * fz) / (m * m);
    }
    block_store<BS>(part, d.partial + b, d.n_blocks, kColsAll);
}

// ---------------------------------------------------------------- O(N^2) FP64 path

// integrator.rs:28-45: kick, scale, drift, wrap into buffer cur ^ 1.
template <int BS>
__global__ __launch_bounds__(BS) void k_drift(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = d.ctrl[G.rep];
    if (C.flags & F_EXIT) return;
    const RepParams& P = d.params[G.rep];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    const bool valid = (s - G.start) < G.count;
    const int cur = C.cur, nx = cur ^ 1;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (valid) {
        const double dt = P.dt, lam = C.lambda, m = G.mass;
        const double c = P.cm[G.type];
        const double* fx = (const double*)d.fx[cur];
        const double* fy = (const double*)d.fy[cur];
        const double* fz = (const double*)d.fz[cur];
        const double vkx = d.vx[cur][s] + fx[s] * c;
        const double vky = d.vy[cur][s] + fy[s] * c;
        const double vkz = d.vz[cur][s] + fz[s] * c;
        part[S_KVX] = vkx * m;
        part[S_KVY] = vky * m;
        part[S_KVZ] = vkz * m;
        part[S_KV2] = m * ((vkx * vkx + vky * vky) + vkz * vkz);
        const double v1x = vkx * lam, v1y = vky * lam, v1z = vkz * lam;
        d.vx[nx][s] = v1x;
        d.vy[nx][s] = v1y;
        d.vz[nx][s] = v1z;
        d.x[nx][s] = wrap1(d.x[cur][s] + v1x * dt, C.box[0]);
        d.y[nx][s] = wrap1(d.y[cur][s] + v1y * dt, C.box[1]);
        d.z[nx][s] = wrap1(d.z[cur][s] + v1z * dt, C.box[2]);
        part[S_VMAX] = (v1x * v1x + v1y * v1y) + v1z * v1z;
    }
    block_store<BS>(part, d.partial + b, d.n_blocks,
                    (P.th_kind == 2 ? kColsKick : 0u) | kColsDrift);
}

// update_force by all pairs: thread per i, j tiles staged through LDS, j
// ascending from the start of i's type to the end of the replica
// (potential.rs:171-177), i == j skipped.  FORCE_ONLY evaluates at x[cur]
// and stores F; otherwise positions come from the drift buffer and the
// second half kick and barostat scale follow (integrator.rs:46-58).
template <int BS, bool MULTI, bool FORCE_ONLY>
__global__ __launch_bounds__(BS) void k_allpairs(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const int r = G.rep;
    const RepCtrl& C = d.ctrl[r];
    if (!FORCE_ONLY && (C.flags & F_EXIT)) return;
    const RepParams& P = d.params[r];
    const RepInfo R = d.reps[r];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    const bool valid = (s - G.start) < G.count;
    const int cur = C.cur;
    const int src = FORCE_ONLY ? cur : (cur ^ 1);
    const double* xs = d.x[src];
    const double* ys = d.y[src];
    const double* zs = d.z[src];
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const int ti = G.type;
    const int T = d.n_types;
    __shared__ double tx[BS], ty[BS], tz[BS];
    __shared__ uint8_t tt[BS];
    const double xi = valid ? xs[s] : 0.0;
    const double yi = valid ? ys[s] : 0.0;
    const double zi = valid ? zs[s] : 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0, vir = 0.0;
    const PairParam p0 = d.pairs[ti * T + ti];
    for (int64_t t0 = G.start; t0 < R.end; t0 += BS) {
        const int64_t jl = t0 + threadIdx.x;
        __syncthreads();
        if (jl < R.end) {
            tx[threadIdx.x] = xs[jl];
            ty[threadIdx.x] = ys[jl];
            tz[threadIdx.x] = zs[jl];
            if (MULTI) tt[threadIdx.x] = d.slot_type[jl];
        }
        __syncthreads();
        const int n = (R.end - t0) < BS ? (int)(R.end - t0) : BS;
        if (valid) {
            for (int k = 0; k < n; ++k) {
                if (t0 + k == s) continue;
                const double dx = min_image(tx[k] - xi, bx);
                const double dy = min_image(ty[k] - yi, by);
                const double dz = min_image(tz[k] - zi, bz);
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                const PairParam& pp = MULTI ? d.pairs[ti * T + tt[k]] : p0;
                if (r2 <= pp.r2max) pair_exact(dx, dy, dz, r2, pp, fx, fy, fz, pot, vir);
            }
        }
    }
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (valid) {
        const double m = G.mass;
        const double c = P.cm[ti];
        double vx, vy, vz;
        if (FORCE_ONLY) {
            vx = d.vx[cur][s];
            vy = d.vy[cur][s];
            vz = d.vz[cur][s];
        } else {
            vx = d.vx[src][s] + fx * c;
            vy = d.vy[src][s] + fy * c;
            vz = d.vz[src][s] + fz * c;
            const double mu = C.myu;
            d.x[cur][s] = xi * mu;
            d.y[cur][s] = yi * mu;
            d.z[cur][s] = zi * mu;
            d.vx[cur][s] = vx;
            d.vy[cur][s] = vy;
            d.vz[cur][s] = vz;
        }
        ((double*)d.fx[cur])[s] = fx;
        ((double*)d.fy[cur])[s] = fy;
        ((double*)d.fz[cur])[s] = fz;
        d.pot[s] = pot;
        d.vir[s] = vir;
        part[S_MVX] = vx * m;
        part[S_MVY] = vy * m;
        part[S_MVZ] = vz * m;
        part[S_MV2] = m * ((vx * vx + vy * vy) + vz * vz);
        part[S_POT] = pot;
        part[S_VIR] = vir;
        part[S_WV] = (vx * vx + vy * vy) + vz * vz;
        part[S_WA] = ((fx * fx + fy * fy) + fz * fz) / (m * m);
    }
    if (FORCE_ONLY) {
        block_store<BS>(part, d.partial + b, d.n_blocks, kColsAll);
    } else {
        block_store<BS>(part, d.partial + b, d.n_blocks, kColsState);
    }
}

// ---------------------------------------------------------------- finalize

__device__ __forceinline__ void thermo(const double* S, const GroupInfo& G, const double* box,
                                       bool kicked, double& temperature, double& pressure) {
    const int ox = kicked ? S_KVX : S_MVX;
    const int o2 = kicked ? S_KV2 : S_MV2;
    const double M = G.sum_m;
    const double cx = S[ox] / M, cy = S[ox + 1] / M, cz = S[ox + 2] / M;
    // sum m |v - v_com|^2 = sum m |v|^2 - M |v_com|^2  (energy.rs:25-37, pressure.rs:11-14)
    const double th2 = S[o2] - M * ((cx * cx + cy * cy) + cz * cz);
    const double thermal = th2 / 2.0;
    // temperature.rs:4-7
    temperature = (2.0 * thermal) / (3.0 * (double)G.count * kKB) * 100.0;
    // pressure.rs:9-19
    const double volume = box[0] * box[1] * box[2];
    pressure = (th2 + (-S[S_VIR]) * 0.5) / volume / 3.0;
}


// Advance a replica's controls after a launch (M_STEP) or prepare the first
// step of a call (M_PRIME).  S holds the group sums of the replica's types.
__device__ void advance(RepCtrl& C, const RepParams& P, const RepInfo& R, const GroupInfo* G,
                        const double* S, int mode) {
    const int T = R.n_types;
    const double dt = P.dt;
    const bool baro = P.ba_kind == 1;
    bool controls = true, decide = true;
    if (mode == M_STEP) {
        C.launches++;
        const uint32_t f = C.flags;
        if (f & F_NULL) {
            // the grid now holds the input positions; the step is still pending
            C.fill_epoch = C.ins_epoch;
            C.flags = (f & F_HAVE_LIST) | F_GRID_FRESH;
            C.nulls++;
            controls = false;
        } else if (f & F_SPLIT_A) {
            // drifted state and its grid are ready; the second half follows
            C.fill_epoch = C.ins_epoch;
            C.flags = F_HAVE_LIST | F_SPLIT_B;
            C.splits++;
            controls = false;
            decide = false;
        } else {
            C.step++;
            C.lambda_used = C.lambda;
            C.myu_used = C.myu;
            if (P.th_kind == 2) {
                // thermostat.rs:59-65: psi advanced with the pre-scaling temperature
                for (int t = 0; t < T; ++t) {
                    double tk, pk;
                    thermo(S + t * NS, G[t], C.box, true, tk, pk);
                    const double psi_dot = -((P.th_target / tk) - 1.0) / P.th_tau;
                    C.psi += psi_dot * (dt / 2.0);
                }
            }
            C.psi_used = C.psi;
            if (baro) {
                // barostat.rs:45 once per type (integrator.rs:54-58 loops over types)
                for (int t = 0; t < T; ++t) {
                    C.box[0] *= C.myu;
                    C.box[1] *= C.myu;
                    C.box[2] *= C.myu;
                }
            }
            uint32_t nf = f & F_HAVE_LIST;
            if (P.list_mode) {
                if (f & (F_REBUILD | F_SPLIT_B)) {
                    C.D = 0.0;
                    C.s = 1.0;
                    C.rebuilds++;
                    nf |= F_HAVE_LIST;
                }
                if (!(f & F_SPLIT_B)) {
                    // the list of a split step was built after this step's drift
                    double vmax = 0.0;
                    for (int t = 0; t < T; ++t) vmax = fmax(vmax, S[t * NS + S_VMAX]);
                    C.D += sqrt(vmax) * (1.0 + 1e-12) * dt / C.s;
                }
                if (baro) C.s *= C.myu;
            }
            if (f & F_INSERT) {
                C.fill_epoch = C.ins_epoch;
                nf |= F_GRID_FRESH;
            }
            // k_step wrote buffer cur ^ 1 (a split step ends in cur)
            if (P.list_mode && !(f & F_SPLIT_B)) C.cur ^= 1;
            C.flags = nf;
        }
    } else {
        C.flags &= (F_HAVE_LIST | F_GRID_FRESH);
        // Nose-Hoover: start from the caller's psi, or from psi after the last
        // completed update when continuing inside one call
        C.psi = P.use_psi_in ? P.psi_in : C.psi_used;
    }
    if (controls) {
        // integrator.rs:16-27: mu and lambda per type in order, last type wins
        double lam = 1.0, mu = 1.0;
        bool bad = false;
        for (int t = 0; t < T; ++t) {
            double tt, pp;
            thermo(S + t * NS, G[t], C.box, false, tt, pp);
            if (baro) {
                const double m3 = 1.0 + dt * P.ba_beta / P.ba_tau * (pp - P.ba_target);
                mu = cbrt(m3);
                if (!(m3 > 0.0)) bad = true;
            }
            if (P.th_kind == 1) {
                const double l2 = 1.0 + dt / P.th_tau * (P.th_target / tt - 1.0);
                lam = sqrt(l2);
            } else if (P.th_kind == 2) {
                const double psi_dot = -((P.th_target / tt) - 1.0) / P.th_tau;
                C.psi += psi_dot * (dt / 2.0);
                lam = exp(-C.psi * dt / 2.0);
            }
        }
        if (!isfinite(lam) || !isfinite(mu)) bad = true;
        C.lambda = lam;
        C.myu = mu;
        if (bad) C.flags |= F_BAD;
    }
    if (P.list_mode && decide) {
        // |v + F dt / 2m| <= |v| + (dt / 2) |F| / m for every particle
        double wmax = 0.0;
        for (int t = 0; t < T; ++t)
            wmax = fmax(wmax, (sqrt(S[t * NS + S_WV]) + (dt / 2.0) * sqrt(S[t * NS + S_WA])) *
                                  (1.0 + 1e-12));
        const double w = C.lambda * wmax * dt / C.s;
        const double thr = P.rlist - P.rc_max / C.s - P.margin;
        bool need = !(C.flags & F_HAVE_LIST) || P.rebuild_every_step || !(2.0 * (C.D + w) < thr);
        if (need) {
            const double w0 = C.lambda * wmax * dt;
            if (!(2.0 * w0 < P.rlist - P.rc_max - P.margin)) {
                // one drift exceeds what any list built before it can cover
                C.flags |= F_SPLIT_A;
                C.ins_epoch = C.fill_epoch + 1u;
                const int gbf = C.ins_epoch & 1;
                for (int k = 0; k < 3; ++k) C.gbox[gbf][k] = C.box[k];
                grid_dims(C.gbox[gbf], P.rlist, R.cell_budget, C.gdim[gbf]);
            } else if (C.flags & F_GRID_FRESH) {
                C.flags |= F_REBUILD;
            } else {
                C.flags |= F_NULL;
                C.ins_epoch = C.fill_epoch + 1u;
                const int gbf = C.ins_epoch & 1;
                for (int k = 0; k < 3; ++k) C.gbox[gbf][k] = C.box[k];
                grid_dims(C.gbox[gbf], P.rlist, R.cell_budget, C.gdim[gbf]);
            }
        }
        if (!(C.flags & (F_NULL | F_SPLIT_A))) {
            const bool rb = (C.flags & F_REBUILD) != 0;
            const double d1 = rb ? 0.0 : C.D;
            const double s1 = rb ? 1.0 : C.s;
            const double w1 = C.lambda * wmax * dt / s1;
            const double s2 = s1 * (baro ? C.myu : 1.0);
            const double thr2 = P.rlist - P.rc_max / s2 - P.margin;
            if (P.rebuild_every_step ||
                !(2.0 * (d1 + w1 * (2.0 + P.insert_growth)) < thr2)) {
                C.flags |= F_INSERT;
                C.ins_epoch = C.fill_epoch + 1u;
                const int gbf = C.ins_epoch & 1;
                double gb[3] = {C.box[0], C.box[1], C.box[2]};
                if (baro)
                    for (int t = 0; t < T; ++t)
                        for (int k = 0; k < 3; ++k) gb[k] *= C.myu;
                for (int k = 0; k < 3; ++k) C.gbox[gbf][k] = gb[k];
                grid_dims(C.gbox[gbf], P.rlist, R.cell_budget, C.gdim[gbf]);
            }
        }
    }
    if (C.step >= P.target_step) {
        C.flags |= F_EXIT;
    } else if (C.step + 1 == P.target_step) {
        C.flags |= F_WRITE_PV;
    }
}

// grid (slices, replicas).  Each block reduces its slice of a group's block
// partials in a fixed order; with several slices the last block to finish
// (release / acquire on a per-replica counter) adds the slice sums in slice
// order, so the result does not depend on which block finishes last.  Then one
// thread advances the replica's controls.
template <int BS>
__global__ __launch_bounds__(BS) void k_finalize(DevState d, int mode, int S) {
    const int sl = blockIdx.x;
    const int r = blockIdx.y;
    const RepInfo R = d.reps[r];
    RepCtrl* Cp = d.ctrl + r;
    const uint32_t f0 = Cp->flags;
    if (mode == M_STEP && (f0 & F_EXIT)) return;
    const bool reduce = mode == M_REDUCE || (mode == M_STEP && !(f0 & (F_NULL | F_SPLIT_A)));
    __shared__ int s_last;
    if (!reduce && sl != 0) return;
    if (reduce) {
        const int64_t nb = d.n_blocks;
        for (int t = 0; t < R.n_types; ++t) {
            const int g = R.first_group + t;
            const GroupInfo G = d.groups[g];
            const int lo = (int)((int64_t)G.nblocks * sl / S);
            const int hi = (int)((int64_t)G.nblocks * (sl + 1) / S);
            double acc[NS];
#pragma unroll
            for (int k = 0; k < NS; ++k) acc[k] = 0.0;
            // column-major partials: consecutive threads read consecutive blocks;
            // four rows per thread are loaded before any is summed
            for (int b0 = lo + threadIdx.x; b0 < hi; b0 += 4 * BS) {
                double v[4][NS];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int bb = b0 + u * BS;
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        v[u][k] = bb < hi ? d.partial[k * nb + G.first_block + bb] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int k = 0; k < NS; ++k)
                        acc[k] = k >= kFirstMax ? fmax(acc[k], v[u][k]) : acc[k] + v[u][k];
            }
            if (S == 1)
                block_store<BS>(acc, d.sums + (int64_t)g * NS, 1, kColsAll);
            else
                block_store<BS>(acc, d.part2 + ((int64_t)g * S + sl) * NS, 1, kColsAll);
        }
        if (S > 1) {
            if (threadIdx.x == 0) {
                __threadfence();
                const unsigned int done = atomicAdd(&d.fin_count[r], 1u);
                s_last = done == (unsigned int)(S - 1);
            }
            __syncthreads();
            if (!s_last) return;
            __threadfence();
            const int k = threadIdx.x;
            if (k < NS) {
                for (int t = 0; t < R.n_types; ++t) {
                    const int64_t g = R.first_group + t;
                    const double* row = d.part2 + g * S * NS;
                    double a = row[k];
                    for (int q = 1; q < S; ++q)
                        a = k >= kFirstMax ? fmax(a, row[q * NS + k]) : a + row[q * NS + k];
                    d.sums[g * NS + k] = a;
                }
            }
            if (threadIdx.x == 0) d.fin_count[r] = 0u;
        }
    }
    if (mode == M_REDUCE) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        RepCtrl C = *Cp;
        advance(C, d.params[r], R, d.groups + R.first_group, d.sums + (int64_t)R.first_group * NS,
                mode);
        *Cp = C;
    }
}

// ---------------------------------------------------------------- small helpers

template <int BS>
__global__ __launch_bounds__(BS) void k_scale(DevState d, int64_t start, int64_t count,
                                              double fac, int which) {
    const int64_t i = start + (int64_t)blockIdx.x * BS + threadIdx.x;
    if (i >= start + count) return;
    const int cur = d.ctrl[0].cur;
    if (which == 0) {
        d.vx[cur][i] *= fac;
        d.vy[cur][i] *= fac;
        d.vz[cur][i] *= fac;
    } else {
        d.x[cur][i] *= fac;
        d.y[cur][i] *= fac;
        d.z[cur][i] *= fac;
    }
}

template <int BS>
__global__ __launch_bounds__(BS) void k_wrap(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = d.ctrl[G.rep];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    if ((s - G.start) >= G.count) return;
    const int cur = C.cur;
    d.x[cur][s] = wrap1(d.x[cur][s], C.box[0]);
    d.y[cur][s] = wrap1(d.y[cur][s], C.box[1]);
    d.z[cur][s] = wrap1(d.z[cur][s], C.box[2]);
}

// ---------------------------------------------------------------- launchers

#define MD_BS_SWITCH(bs, F)                 \
    switch (bs) {                           \
        case 64: F(64); break;              \
        case 128: F(128); break;            \
        case 256: F(256); break;            \
        case 512: F(512); break;            \
        default: return hipErrorInvalidValue; \
    }

hipError_t launch_step(const DevState& d, bool fp32, int bs, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define L(B)                                                                                  \
    if (fp32) {                                                                               \
        if (multi) hipLaunchKernelGGL((k_step<float, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((k_step<float, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    } else {                                                                                  \
        if (multi) hipLaunchKernelGGL((k_step<double, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((k_step<double, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    }
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_force_eval(const DevState& d, bool fp32, int bs, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define L(B)                                                                                        \
    if (fp32) {                                                                                     \
        if (multi) hipLaunchKernelGGL((k_force_eval<float, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((k_force_eval<float, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    } else {                                                                                        \
        if (multi) hipLaunchKernelGGL((k_force_eval<double, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((k_force_eval<double, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    }
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_finalize(const DevState& d, int mode, int n_reps, int slices, hipStream_t st) {
    if (mode == M_PRIME) slices = 1;
    hipLaunchKernelGGL((k_finalize<512>), dim3(slices, n_reps), dim3(512), 0, st, d, mode, slices);
    return hipGetLastError();
}

hipError_t launch_drift(const DevState& d, int bs, hipStream_t st) {
#define L(B) hipLaunchKernelGGL((k_drift<B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_allpairs(const DevState& d, int bs, bool force_only, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define L(B)                                                                                       \
    if (force_only) {                                                                              \
        if (multi) hipLaunchKernelGGL((k_allpairs<B, true, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((k_allpairs<B, false, true>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    } else {                                                                                       \
        if (multi) hipLaunchKernelGGL((k_allpairs<B, true, false>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((k_allpairs<B, false, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    }
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_insert_all(const DevState& d, int bs, hipStream_t st) {
#define L(B) hipLaunchKernelGGL((k_insert_all<B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_reduce_state(const DevState& d, bool fp32, int bs, hipStream_t st) {
#define L(B)                                                                            \
    if (fp32) hipLaunchKernelGGL((k_reduce_state<float, B>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else hipLaunchKernelGGL((k_reduce_state<double, B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_scale(const DevState& d, int bs, int64_t start, int64_t count, double fac,
                        int which, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((count + 255) / 256);
    hipLaunchKernelGGL((k_scale<256>), dim3(nb), dim3(256), 0, st, d, start, count, fac, which);
    (void)bs;
    return hipGetLastError();
}

hipError_t launch_wrap(const DevState& d, int bs, hipStream_t st) {
#define L(B) hipLaunchKernelGGL((k_wrap<B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

}  // namespace md
