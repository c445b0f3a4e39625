// md_kernels.hip -- gfx950 kernels of the MD step engine.
//
// Per time step of a replica the engine runs two launches:
//   k_step      one thread per particle slot: half kick, thermostat scale,
//               drift, wrap (integrator.rs:28-45), Verlet-list force
//               (potential.rs:158-216) with each neighbour's drift recomputed
//               from its input record, second half kick and barostat scale
//               (integrator.rs:47-58), per-block FP64 partial sums, and when
//               asked a fill of the cell grid with the output positions.
//               The same kernel also serves the launches that do not advance
//               time: filling the grid with the input positions (F_NULL),
//               building the Verlet lists brick by brick (F_BUILD), and the
//               two halves of a step whose drift exceeds the skin (F_SPLIT_*).
//   k_finalize  fixed-order reduction of the block partials, then the
//               Berendsen / Nose-Hoover lambda, psi and Berendsen mu of the
//               next step (thermostat.rs:24-72, barostat.rs:21-56), box
//               rescale and the list-validity decision for the next launch.
//               Everything stays on the device; the host replays a graph.
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

// In-place exclusive scan of a[0..n) (n <= kBrickCells); a[n] receives the total.
template <int BS>
__device__ void block_exclusive_scan(int* a, int n) {
    constexpr int NW = BS / 64;
    __shared__ int wsum[NW];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int chunk = (n + BS - 1) / BS;
    const int lo = t * chunk < n ? t * chunk : n;
    const int hi = lo + chunk < n ? lo + chunk : n;
    int local = 0;
    for (int i = lo; i < hi; ++i) local += a[i];
    int incl = local;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int before = 0;
    for (int i = 0; i < w; ++i) before += wsum[i];
    int run = before + incl - local;
    for (int i = lo; i < hi; ++i) {
        const int c = a[i];
        a[i] = run;
        run += c;
    }
    if (t == BS - 1) a[n] = before + incl;
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

// Sorted insertion of j into the list column of slot i (column-major).
__device__ __forceinline__ void list_push(int32_t* __restrict__ col, int64_t N, int n, int cap,
                                          int32_t j) {
    if (n < cap) {
        int p = n;
        while (p > 0 && col[(int64_t)(p - 1) * N] > j) {
            col[(int64_t)p * N] = col[(int64_t)(p - 1) * N];
            --p;
        }
        col[(int64_t)p * N] = j;
    }
}

__device__ __forceinline__ void list_close(const DevState& d, int64_t slot, int n) {
    d.ncount[slot] = n < d.list_cap ? n : d.list_cap;
    if (n > d.list_cap) {
        atomicOr(&d.stats->list_overflow, 1u);
        atomicMax(&d.stats->max_list, n);
    }
}

// A list under construction: the first kListReg entries stay sorted in
// registers (compile-time indexed insertion), longer lists continue in global
// memory.  close() stores the register part and the count.
struct RegList {
    int32_t v[kListReg];
    int n;
    int32_t* __restrict__ col;
    int64_t N;
    int cap;
    __device__ __forceinline__ RegList(int32_t* c, int64_t stride, int capacity)
        : n(0), col(c), N(stride), cap(capacity) {}
    __device__ __forceinline__ void push(int32_t j) {
        if (n < kListReg) {
#pragma unroll
            for (int p = kListReg - 1; p >= 0; --p) {
                if (p <= n) {
                    const bool keep = (p < n) && (v[p] < j);
                    const bool here = p > 0 ? (v[p > 0 ? p - 1 : 0] < j) : true;
                    v[p] = keep ? v[p] : (here ? j : v[p > 0 ? p - 1 : 0]);
                }
            }
        } else {
            if (n == kListReg) {
#pragma unroll
                for (int p = 0; p < kListReg; ++p)
                    if (p < cap) col[(int64_t)p * N] = v[p];
            }
            list_push(col, N, n, cap, j);
        }
        ++n;
    }
    __device__ __forceinline__ void store() {
        if (n <= kListReg) {
#pragma unroll
            for (int p = 0; p < kListReg; ++p)
                if (p < n && p < cap) col[(int64_t)p * N] = v[p];
        }
    }
};

// Verlet list of one particle straight from the grid (fallback when a brick
// does not fit in LDS): every j of the same replica, type >= own type
// (potential.rs:171-180), j != slot, within rlist under the reference
// minimum image, sorted ascending so that sums follow the reference's j order.
template <typename FT, bool MULTI>
__device__ void build_list_global(const DevState& d, const RepCtrl& C, const RepParams& P,
                                  const RepInfo& R, const Rec<FT>* __restrict__ recs,
                                  int64_t slot) {
    const uint32_t ep = C.fill_epoch;
    const int buf = ep & 1;
    const int* gd = C.gdim[buf];
    const double* gb = C.gbox[buf];
    const Rec<FT> me = recs[slot];
    const int ti = (int)me.type;
    const int cx = cell_coord(me.x, gb[0], gd[0]);
    const int cy = cell_coord(me.y, gb[1], gd[1]);
    const int cz = cell_coord(me.z, gb[2], gd[2]);
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double rl2 = P.rlist * P.rlist;
    const int cc = d.cell_cap;
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
                    const Rec<FT> o = recs[j];
                    if (MULTI && (int)o.type < ti) continue;
                    const double dx = min_image(o.x - me.x, bx);
                    const double dy = min_image(o.y - me.y, by);
                    const double dz = min_image(o.z - me.z, bz);
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (!(r2 <= rl2)) continue;
                    list_push(col, d.n_slots, n, d.list_cap, j);
                    ++n;
                }
            }
        }
    }
    list_close(d, slot, n);
}

// Lists of every particle of a replica, one brick of cells at a time: the
// brick and a one-cell halo are staged in LDS (slot ids, positions, types),
// then each particle of the brick scans its 27 neighbour cells there.
// Bricks are shared out among the replica's blocks.
template <typename FT, int BS, bool MULTI>
__device__ void brick_build(const DevState& d, const RepCtrl& C, const RepParams& P,
                            const RepInfo& R, int block_in_rep) {
    const uint32_t ep = C.fill_epoch;
    const int buf = ep & 1;
    const int* gd = C.gdim[buf];
    const int* be = C.bedge[buf];
    const int whole = C.bwhole[buf];
    const Rec<FT>* __restrict__ recs = (const Rec<FT>*)d.rec[C.build_src];
    const unsigned long long* __restrict__ hdr = d.cell_hdr[buf];
    const int32_t* __restrict__ cs = d.cell_slots[buf];
    const int cc = d.cell_cap;
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double rl2 = P.rlist * P.rlist;
    const int64_t N = d.n_slots;
    int nbk[3], hd[3];
    for (int k = 0; k < 3; ++k) {
        const bool w = (whole >> k) & 1;
        nbk[k] = w ? 1 : (gd[k] + be[k] - 1) / be[k];
        hd[k] = w ? gd[k] : be[k] + 2;
    }
    const int nbricks = nbk[0] * nbk[1] * nbk[2];
    const int H = hd[0] * hd[1] * hd[2];
    __shared__ int s_off[kBrickCells + 1];
    __shared__ double s_x[kBrickAtoms], s_y[kBrickAtoms], s_z[kBrickAtoms];
    __shared__ int32_t s_id[kBrickAtoms];
    __shared__ uint16_t s_hc[kBrickAtoms];
    __shared__ uint8_t s_ty[MULTI ? kBrickAtoms : 1];
    for (int bi = block_in_rep; bi < nbricks; bi += R.bld_count) {
        const int bc[3] = {bi / (nbk[1] * nbk[2]), (bi / nbk[2]) % nbk[1], bi % nbk[2]};
        int org[3], ncen[3];
        for (int k = 0; k < 3; ++k) {
            const bool w = (whole >> k) & 1;
            org[k] = w ? 0 : bc[k] * be[k] - 1;
            const int rest = gd[k] - bc[k] * be[k];
            ncen[k] = w ? gd[k] : (rest < be[k] ? rest : be[k]);
        }
        // halo cell h -> grid cell (halo coordinates wrap periodically)
        auto cell_of = [&](int h) -> int64_t {
            const int h0 = h / (hd[1] * hd[2]), h1 = (h / hd[2]) % hd[1], h2 = h % hd[2];
            const int c0 = ((org[0] + h0) % gd[0] + gd[0]) % gd[0];
            const int c1 = ((org[1] + h1) % gd[1] + gd[1]) % gd[1];
            const int c2 = ((org[2] + h2) % gd[2] + gd[2]) % gd[2];
            return R.cell_base + ((int64_t)c0 * gd[1] + c1) * gd[2] + c2;
        };
        for (int h = threadIdx.x; h < H; h += BS) {
            const unsigned long long w = hdr[cell_of(h)];
            int cnt = (uint32_t)(w >> 32) == ep ? (int)(uint32_t)w : 0;
            s_off[h] = cnt < cc ? cnt : cc;
        }
        __syncthreads();
        block_exclusive_scan<BS>(s_off, H);
        const int total = s_off[H];
        if (total > kBrickAtoms) {
            // too dense for LDS: lists of the brick's particles straight from the grid
            if (threadIdx.x == 0) atomicOr(&d.stats->brick_fallback, 1u);
            const int ncell = ncen[0] * ncen[1] * ncen[2];
            for (int q = threadIdx.x; q < ncell; q += BS) {
                const int q0 = q / (ncen[1] * ncen[2]), q1 = (q / ncen[2]) % ncen[1], q2 = q % ncen[2];
                const int h = (((whole & 1) ? q0 : q0 + 1) * hd[1] + ((whole & 2) ? q1 : q1 + 1)) * hd[2] +
                              ((whole & 4) ? q2 : q2 + 1);
                const int64_t cell = cell_of(h);
                const int cnt = s_off[h + 1] - s_off[h];
                for (int k = 0; k < cnt; ++k)
                    build_list_global<FT, MULTI>(d, C, P, R, recs, cs[cell * cc + k]);
            }
            __syncthreads();
            continue;
        }
        for (int h = threadIdx.x; h < H; h += BS) {
            const int64_t cell = cell_of(h);
            const int base = s_off[h], cnt = s_off[h + 1] - base;
            for (int k = 0; k < cnt; ++k) {
                const int32_t j = cs[cell * cc + k];
                const Rec<FT>& o = recs[j];
                s_x[base + k] = o.x;
                s_y[base + k] = o.y;
                s_z[base + k] = o.z;
                s_id[base + k] = j;
                s_hc[base + k] = (uint16_t)h;
                if (MULTI) s_ty[base + k] = (uint8_t)o.type;
            }
        }
        __syncthreads();
        for (int a = threadIdx.x; a < total; a += BS) {
            const int h = s_hc[a];
            const int hc[3] = {h / (hd[1] * hd[2]), (h / hd[2]) % hd[1], h % hd[2]};
            bool center = true;
            int lo[3], hi[3];
            for (int k = 0; k < 3; ++k) {
                if ((whole >> k) & 1) {
                    lo[k] = gd[k] >= 3 ? -1 : 0;
                    hi[k] = gd[k] >= 2 ? 1 : 0;
                } else {
                    center = center && hc[k] >= 1 && hc[k] <= ncen[k];
                    lo[k] = -1;
                    hi[k] = 1;
                }
            }
            if (!center) continue;
            const int32_t i = s_id[a];
            const int ti = MULTI ? (int)s_ty[a] : 0;
            const double xi = s_x[a], yi = s_y[a], zi = s_z[a];
            RegList lst(d.nbr + i, N, d.list_cap);
            for (int o0 = lo[0]; o0 <= hi[0]; ++o0) {
                const int g0 = ((whole & 1) ? (hc[0] + o0 + gd[0]) % gd[0] : hc[0] + o0);
                for (int o1 = lo[1]; o1 <= hi[1]; ++o1) {
                    const int g1 = ((whole & 2) ? (hc[1] + o1 + gd[1]) % gd[1] : hc[1] + o1);
                    for (int o2 = lo[2]; o2 <= hi[2]; ++o2) {
                        const int g2 = ((whole & 4) ? (hc[2] + o2 + gd[2]) % gd[2] : hc[2] + o2);
                        const int h2 = (g0 * hd[1] + g1) * hd[2] + g2;
                        for (int b2 = s_off[h2]; b2 < s_off[h2 + 1]; ++b2) {
                            const int32_t j = s_id[b2];
                            if (j == i) continue;
                            if (MULTI && (int)s_ty[b2] < ti) continue;
                            const double dx = min_image(s_x[b2] - xi, bx);
                            const double dy = min_image(s_y[b2] - yi, by);
                            const double dz = min_image(s_z[b2] - zi, bz);
                            const double r2 = (dx * dx + dy * dy) + dz * dz;
                            if (!(r2 <= rl2)) continue;
                            lst.push(j);
                        }
                    }
                }
            }
            lst.store();
            list_close(d, i, lst.n);
        }
        __syncthreads();
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
    const Rec<FT>* __restrict__ rin = (const Rec<FT>*)d.rec[cur];
    Rec<FT>* __restrict__ rnx = (Rec<FT>*)d.rec[nx];
    double* const part_out = d.partial + b;
    const int64_t pstride = d.n_blocks;
    if (flags & F_NULL) {
        if (valid) grid_insert(d, C, R, s, rin[s].x, rin[s].y, rin[s].z);
        return;
    }
    if (flags & F_BUILD) return;  // k_build's launch
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
            Rec<FT> a = rin[s];
            const double vkx = a.vx + (double)a.fx * c;
            const double vky = a.vy + (double)a.fy * c;
            const double vkz = a.vz + (double)a.fz * c;
            part[S_KVX] = vkx * m;
            part[S_KVY] = vky * m;
            part[S_KVZ] = vkz * m;
            part[S_KV2] = m * ((vkx * vkx + vky * vky) + vkz * vkz);
            a.vx = vkx * lam;
            a.vy = vky * lam;
            a.vz = vkz * lam;
            a.x = wrap1(a.x + a.vx * dt, bx);
            a.y = wrap1(a.y + a.vy * dt, by);
            a.z = wrap1(a.z + a.vz * dt, bz);
            rnx[s] = a;
            part[S_VMAX] = (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz;
            grid_insert(d, C, R, s, a.x, a.y, a.z);
        }
        block_store<BS>(part, part_out, pstride, kick_cols | kColsDrift);
        return;
    }
    // second half of a split step: drifted state in cur ^ 1, results to cur
    const bool split = (flags & F_SPLIT_B) != 0;
    const Rec<FT>* __restrict__ src = split ? rnx : rin;
    Rec<FT>* __restrict__ dst = split ? (Rec<FT>*)d.rec[cur] : rnx;
    if (valid) {
        const Rec<FT> a = src[s];
        double x1, y1, z1, v1x, v1y, v1z;
        if (split) {
            x1 = a.x;
            y1 = a.y;
            z1 = a.z;
            v1x = a.vx;
            v1y = a.vy;
            v1z = a.vz;
        } else {
            // integrator.rs:28-34 first half kick
            const double vkx = a.vx + (double)a.fx * c;
            const double vky = a.vy + (double)a.fy * c;
            const double vkz = a.vz + (double)a.fz * c;
            part[S_KVX] = vkx * m;
            part[S_KVY] = vky * m;
            part[S_KVZ] = vkz * m;
            part[S_KV2] = m * ((vkx * vkx + vky * vky) + vkz * vkz);
            // thermostat.rs:55-57 scale, integrator.rs:40-45 drift + wrap
            v1x = vkx * lam;
            v1y = vky * lam;
            v1z = vkz * lam;
            x1 = wrap1(a.x + v1x * dt, bx);
            y1 = wrap1(a.y + v1y * dt, by);
            z1 = wrap1(a.z + v1z * dt, bz);
            part[S_VMAX] = (v1x * v1x + v1y * v1y) + v1z * v1z;
        }
        // potential.rs:158-216 over the Verlet list at the drifted positions;
        // a neighbour's drift is recomputed from its input record with the
        // same operations its own thread performs, so it is bitwise the same
        Acc<FT> acc;
        const int nn = d.ncount[s];
        const PairParam p0 = d.pairs[0];
        for (int k = 0; k < nn; ++k) {
            const int32_t j = d.nbr[(int64_t)k * d.n_slots + s];
            const Rec<FT> o = src[j];
            const int tj = MULTI ? (int)o.type : 0;
            double xj = o.x, yj = o.y, zj = o.z;
            if (!split) {
                const double cj = MULTI ? P.cm[tj] : c;
                xj = wrap1(xj + ((o.vx + (double)o.fx * cj) * lam) * dt, bx);
                yj = wrap1(yj + ((o.vy + (double)o.fy * cj) * lam) * dt, by);
                zj = wrap1(zj + ((o.vz + (double)o.fz * cj) * lam) * dt, bz);
            }
            const double dx = min_image(xj - x1, bx);
            const double dy = min_image(yj - y1, by);
            const double dz = min_image(zj - z1, bz);
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            const PairParam& pp = MULTI ? d.pairs[ti * d.n_types + tj] : p0;
            if (r2 <= pp.r2max) acc.add(dx, dy, dz, r2, pp);
        }
        // integrator.rs:47-53 second half kick, barostat.rs:45-48 scale
        Rec<FT> out;
        out.vx = v1x + (double)acc.fx * c;
        out.vy = v1y + (double)acc.fy * c;
        out.vz = v1z + (double)acc.fz * c;
        out.x = x1 * mu;
        out.y = y1 * mu;
        out.z = z1 * mu;
        out.fx = acc.fx;
        out.fy = acc.fy;
        out.fz = acc.fz;
        out.type = a.type;
        dst[s] = out;
        if (flags & F_WRITE_PV) {
            d.pot[s] = (double)acc.pot;
            d.vir[s] = (double)acc.vir;
        }
        part[S_MVX] = out.vx * m;
        part[S_MVY] = out.vy * m;
        part[S_MVZ] = out.vz * m;
        part[S_MV2] = m * ((out.vx * out.vx + out.vy * out.vy) + out.vz * out.vz);
        part[S_POT] = (double)acc.pot;
        part[S_VIR] = (double)acc.vir;
        part[S_WV] = (out.vx * out.vx + out.vy * out.vy) + out.vz * out.vz;
        part[S_WA] = (((double)acc.fx * (double)acc.fx + (double)acc.fy * (double)acc.fy) +
                      (double)acc.fz * (double)acc.fz) * P.inv_m2[ti];
        if (flags & F_INSERT) grid_insert(d, C, R, s, out.x, out.y, out.z);
    }
    // a split step stored its kick / drift columns in the first half
    block_store<BS>(part, part_out, pstride, split ? kColsState : (kColsState | kick_cols | kColsDrift));
}

// List build launch: runs before k_step every step and returns at once unless
// its replica is flagged F_BUILD (a separate kernel so that the brick build's
// registers and LDS do not lower the occupancy of k_step).
template <typename FT, int BS, bool MULTI>
__global__ __launch_bounds__(BS) void k_build(DevState d) {
    const int b = blockIdx.x;
    const int r = d.bld_rep[b];
    const RepCtrl& C = d.ctrl[r];
    if ((C.flags & (F_BUILD | F_EXIT)) != F_BUILD) return;
    const RepInfo R = d.reps[r];
    brick_build<FT, BS, MULTI>(d, C, d.params[r], R, b - R.bld_first);
}

// Force evaluation at the current positions through the current lists
// (update_force for the list paths).
template <typename FT, int BS, bool MULTI>
__global__ __launch_bounds__(BS) void k_force_eval(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const int r = G.rep;
    const RepCtrl& C = d.ctrl[r];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    const bool valid = (s - G.start) < G.count;
    Rec<FT>* __restrict__ rc = (Rec<FT>*)d.rec[C.cur];
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double m = G.mass;
    const int ti = G.type;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (valid) {
        Rec<FT> a = rc[s];
        Acc<FT> acc;
        const int nn = d.ncount[s];
        for (int k = 0; k < nn; ++k) {
            const int32_t j = d.nbr[(int64_t)k * d.n_slots + s];
            const Rec<FT> o = rc[j];
            const int tj = MULTI ? (int)o.type : 0;
            const double dx = min_image(o.x - a.x, bx);
            const double dy = min_image(o.y - a.y, by);
            const double dz = min_image(o.z - a.z, bz);
            const double r2 = (dx * dx + dy * dy) + dz * dz;
            const PairParam& pp = d.pairs[ti * d.n_types + tj];
            if (r2 <= pp.r2max) acc.add(dx, dy, dz, r2, pp);
        }
        a.fx = acc.fx;
        a.fy = acc.fy;
        a.fz = acc.fz;
        rc[s] = a;
        d.pot[s] = (double)acc.pot;
        d.vir[s] = (double)acc.vir;
        part[S_MVX] = a.vx * m;
        part[S_MVY] = a.vy * m;
        part[S_MVZ] = a.vz * m;
        part[S_MV2] = m * ((a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz);
        part[S_POT] = (double)acc.pot;
        part[S_VIR] = (double)acc.vir;
        part[S_WV] = (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz;
        part[S_WA] = (((double)acc.fx * (double)acc.fx + (double)acc.fy * (double)acc.fy) +
                      (double)acc.fz * (double)acc.fz) / (m * m);
    }
    block_store<BS>(part, d.partial + b, d.n_blocks, kColsAll);
}

template <typename FT, int BS>
__global__ __launch_bounds__(BS) void k_insert_all(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = d.ctrl[G.rep];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    if ((s - G.start) >= G.count) return;
    const Rec<FT>& a = ((const Rec<FT>*)d.rec[C.cur])[s];
    grid_insert(d, C, d.reps[G.rep], s, a.x, a.y, a.z);
}

// Partial sums of the current state (after an upload or a host-side change).
template <typename FT, int BS>
__global__ __launch_bounds__(BS) void k_reduce_state(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = d.ctrl[G.rep];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    const bool valid = (s - G.start) < G.count;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (valid) {
        const Rec<FT> a = ((const Rec<FT>*)d.rec[C.cur])[s];
        const double m = G.mass;
        part[S_MVX] = a.vx * m;
        part[S_MVY] = a.vy * m;
        part[S_MVZ] = a.vz * m;
        part[S_MV2] = m * ((a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz);
        part[S_POT] = d.pot[s];
        part[S_VIR] = d.vir[s];
        const double fx = (double)a.fx, fy = (double)a.fy, fz = (double)a.fz;
        part[S_WV] = (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz;
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
        Rec<double> a = ((const Rec<double>*)d.rec[cur])[s];
        const double vkx = a.vx + a.fx * c;
        const double vky = a.vy + a.fy * c;
        const double vkz = a.vz + a.fz * c;
        part[S_KVX] = vkx * m;
        part[S_KVY] = vky * m;
        part[S_KVZ] = vkz * m;
        part[S_KV2] = m * ((vkx * vkx + vky * vky) + vkz * vkz);
        a.vx = vkx * lam;
        a.vy = vky * lam;
        a.vz = vkz * lam;
        a.x = wrap1(a.x + a.vx * dt, C.box[0]);
        a.y = wrap1(a.y + a.vy * dt, C.box[1]);
        a.z = wrap1(a.z + a.vz * dt, C.box[2]);
        ((Rec<double>*)d.rec[nx])[s] = a;
        part[S_VMAX] = (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz;
    }
    block_store<BS>(part, d.partial + b, d.n_blocks,
                    (P.th_kind == 2 ? kColsKick : 0u) | kColsDrift);
}

// update_force by all pairs: thread per i, j tiles staged through LDS, j
// ascending from the start of i's type to the end of the replica
// (potential.rs:171-177), i == j skipped.  FORCE_ONLY evaluates at the
// current positions and stores F; otherwise positions come from the drift
// buffer and the second half kick and barostat scale follow
// (integrator.rs:46-58).
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
    const Rec<double>* __restrict__ src = (const Rec<double>*)d.rec[FORCE_ONLY ? cur : (cur ^ 1)];
    Rec<double>* __restrict__ dst = (Rec<double>*)d.rec[cur];
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const int ti = G.type;
    const int T = d.n_types;
    __shared__ double tx[BS], ty[BS], tz[BS];
    __shared__ uint8_t tt[BS];
    Rec<double> a;
    if (valid) a = src[s];
    double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0, vir = 0.0;
    const PairParam p0 = d.pairs[ti * T + ti];
    for (int64_t t0 = G.start; t0 < R.end; t0 += BS) {
        const int64_t jl = t0 + threadIdx.x;
        __syncthreads();
        if (jl < R.end) {
            const Rec<double>& o = src[jl];
            tx[threadIdx.x] = o.x;
            ty[threadIdx.x] = o.y;
            tz[threadIdx.x] = o.z;
            if (MULTI) tt[threadIdx.x] = (uint8_t)o.type;
        }
        __syncthreads();
        const int n = (R.end - t0) < BS ? (int)(R.end - t0) : BS;
        if (valid) {
            for (int k = 0; k < n; ++k) {
                if (t0 + k == s) continue;
                const double dx = min_image(tx[k] - a.x, bx);
                const double dy = min_image(ty[k] - a.y, by);
                const double dz = min_image(tz[k] - a.z, bz);
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
        if (!FORCE_ONLY) {
            a.vx = a.vx + fx * c;
            a.vy = a.vy + fy * c;
            a.vz = a.vz + fz * c;
            const double mu = C.myu;
            a.x = a.x * mu;
            a.y = a.y * mu;
            a.z = a.z * mu;
        }
        a.fx = fx;
        a.fy = fy;
        a.fz = fz;
        dst[s] = a;
        d.pot[s] = pot;
        d.vir[s] = vir;
        part[S_MVX] = a.vx * m;
        part[S_MVY] = a.vy * m;
        part[S_MVZ] = a.vz * m;
        part[S_MV2] = m * ((a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz);
        part[S_POT] = pot;
        part[S_VIR] = vir;
        part[S_WV] = (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz;
        part[S_WA] = ((fx * fx + fy * fy) + fz * fz) / (m * m);
    }
    block_store<BS>(part, d.partial + b, d.n_blocks, FORCE_ONLY ? kColsAll : kColsState);
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

__device__ void set_fill(RepCtrl& C, const RepParams& P, const RepInfo& R, const double* gb) {
    C.ins_epoch = C.fill_epoch + 1u;
    const int g = C.ins_epoch & 1;
    for (int k = 0; k < 3; ++k) C.gbox[g][k] = gb[k];
    grid_dims(C.gbox[g], P.rlist, R.cell_budget, C.gdim[g]);
    brick_dims(C.gdim[g], R.n_atoms, C.bedge[g], &C.bwhole[g]);
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
            // drifted state and its grid are ready: build lists there, then the second half
            C.fill_epoch = C.ins_epoch;
            C.build_src = C.cur ^ 1;
            C.flags = F_BUILD | F_SPLIT_PENDING;
            C.splits++;
            controls = false;
            decide = false;
        } else if (f & F_BUILD) {
            // lists now describe the build positions
            C.D = 0.0;
            C.s = 1.0;
            C.rebuilds++;
            const uint32_t keep = f & F_GRID_FRESH;
            if (f & F_SPLIT_PENDING) {
                C.flags = F_HAVE_LIST | F_SPLIT_B;
                decide = false;
            } else {
                C.flags = F_HAVE_LIST | keep;
            }
            controls = false;
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
        const bool need =
            !(C.flags & F_HAVE_LIST) || P.rebuild_every_step || !(2.0 * (C.D + w) < thr);
        if (need) {
            const double w0 = C.lambda * wmax * dt;
            if (!(2.0 * w0 < P.rlist - P.rc_max - P.margin)) {
                // one drift exceeds what any list built before it can cover
                C.flags |= F_SPLIT_A;
                set_fill(C, P, R, C.box);
            } else if (C.flags & F_GRID_FRESH) {
                C.flags |= F_BUILD;
                C.build_src = C.cur;
            } else {
                C.flags |= F_NULL;
                set_fill(C, P, R, C.box);
            }
        } else {
            // fill the grid during this step when the list may expire within
            // the next two, so that the rebuild needs no separate fill
            const double s2 = C.s * (baro ? C.myu : 1.0);
            const double thr2 = P.rlist - P.rc_max / s2 - P.margin;
            if (P.rebuild_every_step || !(2.0 * (C.D + w * (2.0 + P.insert_growth)) < thr2)) {
                C.flags |= F_INSERT;
                double gb[3] = {C.box[0], C.box[1], C.box[2]};
                if (baro)
                    for (int t = 0; t < T; ++t)
                        for (int k = 0; k < 3; ++k) gb[k] *= C.myu;
                set_fill(C, P, R, gb);
            }
        }
    }
    if (C.step >= P.target_step) {
        C.flags |= F_EXIT;
    } else if (C.step + 1 == P.target_step) {
        C.flags |= F_WRITE_PV;
    }
}

// Which launches produced block partials to reduce.
__device__ __forceinline__ bool finalize_reduces(int mode, uint32_t f0) {
    return mode == M_REDUCE || (mode == M_STEP && !(f0 & (F_EXIT | F_NULL | F_SPLIT_A | F_BUILD)));
}

// Sum rows [lo, hi) of a group's column-major block partials in a fixed order
// (thread-strided, two rows in flight per thread, then the block tree).
template <int BS>
__device__ void reduce_rows(const DevState& d, const GroupInfo& G, int lo, int hi, double* out) {
    const int64_t nb = d.n_blocks;
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;
    for (int b0 = lo + threadIdx.x; b0 < hi; b0 += 2 * BS) {
        double v[2][NS];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int bb = b0 + u * BS;
#pragma unroll
            for (int k = 0; k < NS; ++k)
                v[u][k] = bb < hi ? d.partial[k * nb + G.first_block + bb] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int k = 0; k < NS; ++k)
                acc[k] = k >= kFirstMax ? fmax(acc[k], v[u][k]) : acc[k] + v[u][k];
    }
    block_store<BS>(acc, out, 1, kColsAll);
}

// First level of a sliced reduction: grid (slices, replicas); slice sums go
// to part2 and the finalize adds them in slice order (the kernel boundary
// orders the two, so no fences are needed).
template <int BS>
__global__ __launch_bounds__(BS) void k_reduce_slices(DevState d, int mode, int S) {
    const int sl = blockIdx.x;
    const int r = blockIdx.y;
    const RepInfo R = d.reps[r];
    if (!finalize_reduces(mode, d.ctrl[r].flags)) return;
    for (int t = 0; t < R.n_types; ++t) {
        const int g = R.first_group + t;
        const GroupInfo G = d.groups[g];
        const int lo = (int)((int64_t)G.nblocks * sl / S);
        const int hi = (int)((int64_t)G.nblocks * (sl + 1) / S);
        reduce_rows<BS>(d, G, lo, hi, d.part2 + ((int64_t)g * S + sl) * NS);
    }
}

// grid (replicas).  Group sums (directly from the block partials, or from the
// slice sums of k_reduce_slices), then one thread advances the replica's
// controls from LDS-staged inputs.
template <int BS>
__global__ __launch_bounds__(BS) void k_finalize(DevState d, int mode, int S) {
    const int r = blockIdx.x;
    const RepInfo R = d.reps[r];
    RepCtrl* Cp = d.ctrl + r;
    const uint32_t f0 = Cp->flags;
    if (mode == M_STEP && (f0 & F_EXIT)) return;
    if (finalize_reduces(mode, f0)) {
        for (int t = 0; t < R.n_types; ++t) {
            const int g = R.first_group + t;
            if (S == 1) {
                const GroupInfo G = d.groups[g];
                reduce_rows<BS>(d, G, 0, G.nblocks, d.sums + (int64_t)g * NS);
            } else {
                const int k = threadIdx.x;
                if (k < NS) {
                    const double* row = d.part2 + (int64_t)g * S * NS;
                    double a = row[k];
                    for (int q = 1; q < S; ++q)
                        a = k >= kFirstMax ? fmax(a, row[q * NS + k]) : a + row[q * NS + k];
                    d.sums[(int64_t)g * NS + k] = a;
                }
            }
        }
    }
    if (mode == M_REDUCE) return;
    // Stage everything the control update reads into LDS with one round of
    // parallel loads: the update itself is a serial chain on one lane and
    // would otherwise wait on each global load in turn.
    __shared__ RepCtrl s_ctrl;
    __shared__ RepParams s_par;
    __shared__ GroupInfo s_grp[kMaxTypes];
    __shared__ double s_sums[kMaxTypes * NS];
    __syncthreads();
    {
        const int T = R.n_types;
        const int nc = (int)(sizeof(RepCtrl) / 4), np = (int)(sizeof(RepParams) / 4);
        const int ng = (int)(sizeof(GroupInfo) / 4) * T;
        const int tid = threadIdx.x;
        if (tid < nc) ((uint32_t*)&s_ctrl)[tid] = ((const uint32_t*)Cp)[tid];
        if (tid >= 128 && tid - 128 < np)
            ((uint32_t*)&s_par)[tid - 128] = ((const uint32_t*)(d.params + r))[tid - 128];
        if (tid >= 256 && tid - 256 < ng)
            ((uint32_t*)s_grp)[tid - 256] = ((const uint32_t*)(d.groups + R.first_group))[tid - 256];
        if (tid >= 384 && tid - 384 < T * NS)
            s_sums[tid - 384] = d.sums[(int64_t)R.first_group * NS + (tid - 384)];
    }
    __syncthreads();
    if (threadIdx.x == 0) advance(s_ctrl, s_par, R, s_grp, s_sums, mode);
    __syncthreads();
    {
        const int nc = (int)(sizeof(RepCtrl) / 4);
        if ((int)threadIdx.x < nc) ((uint32_t*)Cp)[threadIdx.x] = ((const uint32_t*)&s_ctrl)[threadIdx.x];
    }
}

// ---------------------------------------------------------------- small helpers

template <typename FT>
__global__ __launch_bounds__(256) void k_scale(DevState d, int64_t start, int64_t count,
                                               double fac, int which) {
    const int64_t i = start + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= start + count) return;
    Rec<FT>& a = ((Rec<FT>*)d.rec[d.ctrl[0].cur])[i];
    if (which == 0) {
        a.vx *= fac;
        a.vy *= fac;
        a.vz *= fac;
    } else {
        a.x *= fac;
        a.y *= fac;
        a.z *= fac;
    }
}

template <typename FT, int BS>
__global__ __launch_bounds__(BS) void k_wrap(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = d.ctrl[G.rep];
    const int64_t s = (int64_t)b * BS + threadIdx.x;
    if ((s - G.start) >= G.count) return;
    Rec<FT>& a = ((Rec<FT>*)d.rec[C.cur])[s];
    a.x = wrap1(a.x, C.box[0]);
    a.y = wrap1(a.y, C.box[1]);
    a.z = wrap1(a.z, C.box[2]);
}

// ---------------------------------------------------------------- launchers

#define MD_BS_SWITCH(bs, F)                   \
    switch (bs) {                             \
        case 64: F(64); break;                \
        case 128: F(128); break;              \
        case 256: F(256); break;              \
        case 512: F(512); break;              \
        default: return hipErrorInvalidValue; \
    }

#define MD_LAUNCH_FT_MULTI_G(K, B, G)                                                     \
    if (fp32) {                                                                            \
        if (multi) hipLaunchKernelGGL((K<float, B, true>), dim3(G), dim3(B), 0, st, d);      \
        else hipLaunchKernelGGL((K<float, B, false>), dim3(G), dim3(B), 0, st, d);           \
    } else {                                                                               \
        if (multi) hipLaunchKernelGGL((K<double, B, true>), dim3(G), dim3(B), 0, st, d);     \
        else hipLaunchKernelGGL((K<double, B, false>), dim3(G), dim3(B), 0, st, d);          \
    }

#define MD_LAUNCH_FT_MULTI(K, B)                                                               \
    if (fp32) {                                                                                \
        if (multi) hipLaunchKernelGGL((K<float, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((K<float, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    } else {                                                                                   \
        if (multi) hipLaunchKernelGGL((K<double, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((K<double, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    }

hipError_t launch_step(const DevState& d, bool fp32, int bs, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define L(B) MD_LAUNCH_FT_MULTI(k_step, B)
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_build(const DevState& d, bool fp32, int bs, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define L(B) MD_LAUNCH_FT_MULTI_G(k_build, B, d.n_bld)
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_force_eval(const DevState& d, bool fp32, int bs, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define L(B) MD_LAUNCH_FT_MULTI(k_force_eval, B)
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_finalize(const DevState& d, int mode, int n_reps, int slices, hipStream_t st) {
    if (mode == M_PRIME) slices = 1;
    if (slices > 1)
        hipLaunchKernelGGL((k_reduce_slices<512>), dim3(slices, n_reps), dim3(512), 0, st, d, mode, slices);
    hipLaunchKernelGGL((k_finalize<512>), dim3(n_reps), dim3(512), 0, st, d, mode, slices);
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

hipError_t launch_insert_all(const DevState& d, bool fp32, int bs, hipStream_t st) {
#define L(B)                                                                               \
    if (fp32) hipLaunchKernelGGL((k_insert_all<float, B>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else hipLaunchKernelGGL((k_insert_all<double, B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_reduce_state(const DevState& d, bool fp32, int bs, hipStream_t st) {
#define L(B)                                                                                 \
    if (fp32) hipLaunchKernelGGL((k_reduce_state<float, B>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else hipLaunchKernelGGL((k_reduce_state<double, B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_scale(const DevState& d, bool fp32, int64_t start, int64_t count, double fac,
                        int which, hipStream_t st) {
    if (count <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((count + 255) / 256);
    if (fp32) hipLaunchKernelGGL((k_scale<float>), dim3(nb), dim3(256), 0, st, d, start, count, fac, which);
    else hipLaunchKernelGGL((k_scale<double>), dim3(nb), dim3(256), 0, st, d, start, count, fac, which);
    return hipGetLastError();
}

hipError_t launch_wrap(const DevState& d, bool fp32, int bs, hipStream_t st) {
#define L(B)                                                                         \
    if (fp32) hipLaunchKernelGGL((k_wrap<float, B>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else hipLaunchKernelGGL((k_wrap<double, B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

}  // namespace md
