// md_kernels.hip -- gfx950 kernels of the MD step engine.
//
// One launch per time step of every replica, two kernels:
//   k_step  one thread per particle slot of each of the block's tiles: half
//           kick, thermostat scale, drift, wrap (integrator.rs:28-45),
//           Verlet-list force (potential.rs:158-216) with each neighbour's
//           drift recomputed from its input record, second half kick and
//           barostat scale (integrator.rs:47-58), per-block FP64 partial
//           sums, and when asked a fill of the cell grid with the output
//           positions.  It also serves the launches that do not advance time:
//           filling the grid with the input positions (F_NULL) and the first
//           half of a step whose drift exceeds the skin (F_SPLIT_A).
//   k_fin   per replica, the fixed-order reduction of the block partials,
//           then the Berendsen / Nose-Hoover lambda, psi and Berendsen mu of
//           the next step (thermostat.rs:24-72, barostat.rs:21-56), box
//           rescale and the list-validity decision for the next launch; its
//           other blocks build the Verlet lists brick by brick in launches
//           flagged F_BUILD (k_step does nothing in those).
// Everything stays on the device; the host replays a graph of launches.
// The O(N^2) FP64 path (k_drift + k_allpairs) restates update_force exactly:
// same pair criterion, same operation order, j ascending.
//
// All code is built with -ffp-contract=off: the FP64 paths must not form FMAs
// the reference (Rust) never forms.  The FP32 pair kernel uses explicit fmaf.
#include <math.h>

#include <type_traits>

#include "md_common.h"

namespace md {

// Type of the particle in slot j: full records carry it, step records do not.
template <typename FT>
__device__ __forceinline__ int rec_type(const DevState&, const Rec<FT>* r, int64_t j) {
    return (int)r[j].type;
}
__device__ __forceinline__ int rec_type(const DevState& d, const CRec*, int64_t j) { return (int)d.slot_type[j]; }

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
    } else if (4 * n > 3 * d.list_cap) {
        atomicMax(&d.stats->max_list, n);
    }
}

// A list under construction.  Without an LDS column, the first kListReg
// entries stay sorted in registers (compile-time indexed insertion) and a
// longer list continues in global memory (sorted insertion); store() writes
// the register part.  With SEG and a column the builder hands over (room for
// the list capacity, at a stride of seg_stride; dense systems), the scan
// lanes append the staging index of every candidate the FP32 test accepts
// (brick_build), and one lane then runs the FP64 test, compacts and sorts
// the column once (store_seg), in loops every lane runs at the same time:
// in a dense brick some lane of the wave accepts a candidate at nearly every
// step of the scan, and work per accepted entry inside the scan would make
// the whole wave execute it each time.
template <bool SEG>
struct RegList {
    int32_t v[kListReg];
    int n;
    int32_t* __restrict__ col;
    int64_t N;
    int cap;
    int32_t* seg = nullptr;  // LDS column of cap entries at a stride of seg_stride (SEG)
    int seg_stride = 0;
    __device__ __forceinline__ RegList(int32_t* c, int64_t stride, int capacity)
        : n(0), col(c), N(stride), cap(capacity) {}
    // SEG: sort the column's first m entries (slot ids) and store them
    __device__ __forceinline__ void store_seg(int m) {
        for (int q = 1; q < m; ++q) {
            const int32_t x = seg[q * seg_stride];
            int p = q;
            while (p > 0 && seg[(p - 1) * seg_stride] > x) {
                seg[p * seg_stride] = seg[(p - 1) * seg_stride];
                --p;
            }
            seg[p * seg_stride] = x;
        }
        for (int p = 0; p < m; ++p) col[(int64_t)p * N] = seg[p * seg_stride];
    }
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
template <typename RT, bool MULTI>
__device__ void build_list_global(const DevState& d, const RepCtrl& C, const RepParams& P,
                                  const RepInfo& R, const RT* __restrict__ recs,
                                  int64_t slot) {
    const uint32_t ep = C.fill_epoch;
    const int buf = ep & 1;
    const int* gd = C.gdim[buf];
    const double* gb = C.gbox[buf];
    const RT me = recs[slot];
    const int ti = MULTI ? rec_type(d, recs, slot) : 0;
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
                    const int32_t j = cs[(int64_t)k * d.cell_plane + cell];
                    if ((int64_t)j == slot) continue;
                    const RT o = recs[j];
                    if (MULTI && rec_type(d, recs, j) < ti) continue;
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


// The reference's minimum image (potential.rs:186-200) as selects, for the
// list builder's exact tests: the same values as min_image, without a branch
// per component in the candidate loops.
__device__ __forceinline__ double min_image_sel(double r, double b) {
    const double h = b / 2.0;
    return r < -h ? r + b : (r > h ? r - b : r);
}

// LDS of a brick: one object per kernel, shared by the builders of both
// record types.
struct BrickLds {
    int off[kBrickCells + 1];  // staged particles before halo cell h
    int cen[kBrickCells + 1];  // the brick's own (centre) particles before cell h
    double x[kBrickAtoms], y[kBrickAtoms], z[kBrickAtoms];
    float4 pf[kBrickAtoms];  // FP32 copy of the position (w unused): one 16-byte read per candidate
    int32_t id[kBrickAtoms];
    uint16_t cidx[kBrickAtoms];  // centre particle k -> its staging index
    uint16_t hc[kBrickAtoms];    // staged particle -> its halo cell
    uint8_t ty[kBrickAtoms];
    int32_t seg[kBrickSeg];  // long lists of the centre particles, one column each
};
__device__ __forceinline__ BrickLds& brick_lds() {
    __shared__ BrickLds l;
    return l;
}

// Lists of every particle of a replica, one brick of cells at a time: the
// brick and a one-cell halo are staged in LDS (slot ids, positions, types),
// then each particle of the brick scans its 27 neighbour cells there, as nine
// runs of consecutive staged particles (the cells of a z row).  A
// conservative FP32 distance test rejects most candidates before the FP64 test
// of the reference's minimum image decides.
// Bricks are shared out among the replica's blocks.  The barriers order LDS
// only: a brick's list stores stay in flight while the next brick loads.
// Bricks of the grid a replica's lists are built from.
__device__ __forceinline__ int brick_count(const RepCtrl& C) {
    const int buf = C.fill_epoch & 1;
    int n = 1;
    for (int k = 0; k < 3; ++k)
        n *= ((C.bwhole[buf] >> k) & 1) ? 1 : (C.gdim[buf][k] + C.bedge[buf][k] - 1) / C.bedge[buf][k];
    return n;
}

// Bricks bi0, bi0 + stride, ... of a replica (all threads of the block).
// SEG: lists longer than kListReg continue in LDS columns (dense systems).
template <typename RT, int BS, bool MULTI, bool SEG>
__device__ __forceinline__ void brick_build(const DevState& d, const RepCtrl& C, const RepParams& P,
                            const RepInfo& R, const RT* __restrict__ recs, int bi0, int stride) {
    const uint32_t ep = C.fill_epoch;
    const int buf = ep & 1;
    const int* gd = C.gdim[buf];
    const int* be = C.bedge[buf];
    const int whole = C.bwhole[buf];
    const unsigned long long* __restrict__ hdr = d.cell_hdr[buf];
    const int32_t* __restrict__ cs = d.cell_slots[buf];
    const int cc = d.cell_cap;
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double rl2 = P.rlist * P.rlist;
    // FP32 prefilter: positions below 1.01 L_max carry an FP32 error below
    // 2^-24 (1.01 L_max) each, so a minimum-image component in FP32 is within
    // eps = 5.05 * 2^-24 L_max of the FP64 one (two roundings of the
    // positions, the difference, the box and the shift), and |r| within
    // 2 sqrt(3) eps (a branch taken the other way near L/2 doubles it).  The
    // squared sum adds 3 roundings; 2^-20 covers them.
    const float bxf = (float)bx, byf = (float)by, bzf = (float)bz;
    // With every axis tiled into at least four cells, the FP32 copies hold
    // each staged particle's periodic image next to its halo cell, relative
    // to the region's first cell (|coordinate| <= 1.25 L): two particles of
    // adjacent cells (|unwrapped difference| < 2 cells) within r_list
    // (<= 1 cell) then have their minimum-image difference as the plain
    // difference, since the two differ by a multiple of L smaller than
    // 3 cells <= 0.75 L.  The prefilter needs no minimum image, and the
    // roundings stay inside eps below (1.25 + 1.25 + 0.5 < 5.05).
    const bool unwrap = whole == 0 && gd[0] >= 4 && gd[1] >= 4 && gd[2] >= 4;
    const double eps = 5.05 * 0x1p-24 * fmax(bx, fmax(by, bz));
    const double rlm = P.rlist + 2.0 * 1.7320508075688772 * eps;
    const float rl2f = __double2float_ru(rlm * rlm * (1.0 + 0x1p-20));
    const int64_t N = d.n_slots;
    int nbk[3];
    for (int k = 0; k < 3; ++k) {
        const bool w = (whole >> k) & 1;
        nbk[k] = w ? 1 : (gd[k] + be[k] - 1) / be[k];
    }
    const int nbricks = nbk[0] * nbk[1] * nbk[2];
    BrickLds& L = brick_lds();
    int* const s_off = L.off;
    int* const s_cen = L.cen;
    uint16_t* const s_cidx = L.cidx;
    double* const s_x = L.x;
    double* const s_y = L.y;
    double* const s_z = L.z;
    float4* const s_pf = L.pf;
    int32_t* const s_id = L.id;
    uint16_t* const s_hc = L.hc;
    uint8_t* const s_ty = L.ty;
    for (int64_t b64 = bi0; b64 < nbricks; b64 += stride) {  // 64-bit: a stride may be INT_MAX
        const int bi = (int)b64;
        const int bc[3] = {bi / (nbk[1] * nbk[2]), (bi / nbk[2]) % nbk[1], bi % nbk[2]};
        // The regions built from LDS: the brick; when its halo holds more
        // particles than LDS stages (a density fluctuation above the grid's
        // mean), its centre split into sub-bricks along the tiled axes, and a
        // sub-brick that still overflows cell by cell.  A single cell whose
        // halo overflows builds its particles' lists straight from the grid,
        // one particle per thread.  All region bookkeeping is uniform over
        // the block (it follows the LDS totals).
        int borg[3], bcen[3];  // the brick: halo origin (grid cells) and centre extent
        for (int k = 0; k < 3; ++k) {
            const bool w = (whole >> k) & 1;
            borg[k] = w ? 0 : bc[k] * be[k] - 1;
            const int rest = gd[k] - bc[k] * be[k];
            bcen[k] = w ? gd[k] : (rest < be[k] ? rest : be[k]);
        }
        int level = 0;                             // 0 brick, 1 sub-brick, 2 cell
        int sm[3] = {1, 1, 1}, sp[3] = {1, 1, 1};  // sub-brick centre extent, sub-bricks per axis
        int j1 = 0, n1 = 1, j2 = 0, n2 = 1;        // sub-brick / cell index and count
        int sorg[3] = {0, 0, 0}, scen[3] = {1, 1, 1};  // the current sub-brick
        for (;;) {
        int org[3], ncen[3], hd[3];
        for (int k = 0; k < 3; ++k) {
            const bool w = (whole >> k) & 1;
            if (w || level == 0) {
                org[k] = borg[k];
                ncen[k] = bcen[k];
            } else if (level == 1) {
                const int q = k == 0 ? j1 / (sp[1] * sp[2]) : (k == 1 ? (j1 / sp[2]) % sp[1] : j1 % sp[2]);
                org[k] = borg[k] + q * sm[k];
                const int rest = bcen[k] - q * sm[k];
                ncen[k] = rest < sm[k] ? rest : sm[k];
            } else {
                const int e1 = ((whole >> 1) & 1) ? 1 : scen[1], e2 = ((whole >> 2) & 1) ? 1 : scen[2];
                const int q = k == 0 ? j2 / (e1 * e2) : (k == 1 ? (j2 / e2) % e1 : j2 % e2);
                org[k] = sorg[k] + q - 1;
                ncen[k] = 1;
            }
            hd[k] = w ? gd[k] : ncen[k] + 2;
        }
        const int H = hd[0] * hd[1] * hd[2];
        // halo cell h = (h0 hd1 + h1) hd2 + h2, split with float reciprocals
        // (no integer division): for h < 2^13 the estimate is off by at most
        // one, which the remainder corrects
        const int hd12 = hd[1] * hd[2];
        const float rhd12 = 1.0f / (float)hd12, rhd2 = 1.0f / (float)hd[2];
        auto divmod = [](int n, int dv, float rdv, int& q) -> int {
            int qq = (int)((float)n * rdv);
            int rr = n - qq * dv;
            if (rr < 0) {
                --qq;
                rr += dv;
            } else if (rr >= dv) {
                ++qq;
                rr -= dv;
            }
            q = qq;
            return rr;
        };
        auto split = [&](int h, int& h0, int& h1, int& h2) {
            const int rem = divmod(h, hd12, rhd12, h0);
            h2 = divmod(rem, hd[2], rhd2, h1);
        };
        // a halo coordinate's grid cell along axis k (org + h in [-1, 2 gd))
        auto wrapc = [&](int g, int k) -> int {
            g += g < 0 ? gd[k] : 0;
            return g >= gd[k] ? g - gd[k] : g;
        };
        // halo cell h -> grid cell (halo coordinates wrap periodically)
        auto cell_of = [&](int h) -> int64_t {
            int h0, h1, h2;
            split(h, h0, h1, h2);
            const int c0 = wrapc(org[0] + h0, 0), c1 = wrapc(org[1] + h1, 1), c2 = wrapc(org[2] + h2, 2);
            return R.cell_base + ((int64_t)c0 * gd[1] + c1) * gd[2] + c2;
        };
        // the unwrapping shift of axis k for halo coordinate hk (unwrap): the
        // image of the cell's particles next to the region, less the
        // region's first cell
        auto shift = [&](int hk, int k, double b) -> double {
            const int g = org[k] + hk;
            const int w = g < 0 ? -1 : (g >= gd[k] ? 1 : 0);
            return (double)w * b - (double)org[k] * (b / (double)gd[k]);
        };
        // all header loads of a thread in flight at once
        constexpr int HP = (kBrickCells + BS - 1) / BS;
        unsigned long long hw[HP];
#pragma unroll
        for (int q = 0; q < HP; ++q) {
            const int h = threadIdx.x + q * BS;
            hw[q] = h < H ? hdr[cell_of(h)] : 0ull;
        }
#pragma unroll
        for (int q = 0; q < HP; ++q) {
            const int h = threadIdx.x + q * BS;
            if (h < H) {
                const int cnt = (uint32_t)(hw[q] >> 32) == ep ? (int)(uint32_t)hw[q] : 0;
                int h0, h1, h2;
                split(h, h0, h1, h2);
                const bool own = ((whole & 1) || (h0 >= 1 && h0 <= ncen[0])) &&
                                 ((whole & 2) || (h1 >= 1 && h1 <= ncen[1])) &&
                                 ((whole & 4) || (h2 >= 1 && h2 <= ncen[2]));
                s_off[h] = cnt < cc ? cnt : cc;
                s_cen[h] = own ? (cnt < cc ? cnt : cc) : 0;
            }
        }
        lds_barrier();
        block_exclusive_scan2<BS>(s_off, s_cen, H);
        const int total = s_off[H];
        const int n_own = s_cen[H];
        if (total > kBrickAtoms) {
            // the centre cells along tiled axes that can still be split
            const bool split = level < 2 && (level == 0 || ((~whole & 1) && ncen[0] > 1) ||
                                             ((~whole & 2) && ncen[1] > 1) || ((~whole & 4) && ncen[2] > 1));
            if (split && level == 0) {
                // sub-brick centres: halve the longest tiled extent until the
                // halo's expected count at this brick's density fits
                // (fixed indices only: the extents stay in registers)
                const double per = (double)total / (double)H;
                const bool w0 = whole & 1, w1 = (whole >> 1) & 1, w2 = (whole >> 2) & 1;
                int m0 = w0 ? gd[0] : bcen[0], m1 = w1 ? gd[1] : bcen[1], m2 = w2 ? gd[2] : bcen[2];
                for (int it = 0; it < 32; ++it) {
                    const double hv = (double)(w0 ? gd[0] : m0 + 2) * (double)(w1 ? gd[1] : m1 + 2) *
                                      (double)(w2 ? gd[2] : m2 + 2);
                    const int a0 = w0 ? 1 : m0, a1 = w1 ? 1 : m1, a2 = w2 ? 1 : m2;
                    if ((a0 <= 1 && a1 <= 1 && a2 <= 1) || hv * per <= 0.75 * (double)kBrickAtoms) break;
                    if (a0 >= a1 && a0 >= a2) m0 = (m0 + 1) / 2;
                    else if (a1 >= a2) m1 = (m1 + 1) / 2;
                    else m2 = (m2 + 1) / 2;
                }
                sm[0] = m0;
                sm[1] = m1;
                sm[2] = m2;
                n1 = 1;
                for (int k = 0; k < 3; ++k) {
                    sp[k] = ((whole >> k) & 1) ? 1 : (bcen[k] + sm[k] - 1) / sm[k];
                    n1 *= sp[k];
                }
                level = 1;
                j1 = 0;
                lds_barrier();
                continue;
            }
            if (split && level == 1) {
                for (int k = 0; k < 3; ++k) {
                    sorg[k] = org[k] + 1;  // the centre's first cell (tiled axes)
                    scen[k] = ncen[k];
                }
                n2 = 1;
                for (int k = 0; k < 3; ++k)
                    if (!((whole >> k) & 1)) n2 *= scen[k];
                level = 2;
                j2 = 0;
                lds_barrier();
                continue;
            }
            // a region that cannot be split further: its centre particles'
            // lists straight from the grid, one particle per thread (the
            // centre prefix sums map a particle to its halo cell)
            if (threadIdx.x == 0) atomicOr(&d.stats->brick_fallback, 1u);
            for (int k = threadIdx.x; k < n_own; k += BS) {
                int lo = 0, hi = H;  // the h with s_cen[h] <= k < s_cen[h + 1]
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_cen[mid] <= k) lo = mid;
                    else hi = mid;
                }
                build_list_global<RT, MULTI>(d, C, P, R, recs,
                                               cs[(int64_t)(k - s_cen[lo]) * d.cell_plane + cell_of(lo)]);
            }
        } else {
        // staged particle -> its halo cell, centre particle -> staging index
        // (LDS only), then one particle per thread: slot id and position
        // loads of all particles go out together
        for (int h = threadIdx.x; h < H; h += BS) {
            const int a0 = s_off[h], a1 = s_off[h + 1], k0 = s_cen[h];
            const bool own = s_cen[h + 1] > k0;
            for (int a = a0; a < a1; ++a) {
                s_hc[a] = (uint16_t)h;
                if (own) s_cidx[k0 + (a - a0)] = (uint16_t)a;
            }
        }
        lds_barrier();
        for (int a = threadIdx.x; a < total; a += BS) {
            const int h = s_hc[a];
            const int32_t j = cs[(int64_t)(a - s_off[h]) * d.cell_plane + cell_of(h)];
            const RT& o = recs[j];
            s_x[a] = o.x;
            s_y[a] = o.y;
            s_z[a] = o.z;
            if (unwrap) {
                int h0, h1, h2;
                split(h, h0, h1, h2);
                s_pf[a] = make_float4((float)(o.x + shift(h0, 0, bx)), (float)(o.y + shift(h1, 1, by)),
                                      (float)(o.z + shift(h2, 2, bz)), 0.0f);
            } else {
                s_pf[a] = make_float4((float)o.x, (float)o.y, (float)o.z, 0.0f);
            }
            s_id[a] = j;
            if (MULTI) s_ty[a] = (uint8_t)rec_type(d, recs, j);
        }
        lds_barrier();
        // LDS list columns (dense instances): every centre particle of a pass
        // owns a column of list_cap entries in the pool, so a long capacity
        // splits the centre particles into passes instead of sending their
        // lists to global memory
        const int dl = d.list_cap > 1 ? d.list_cap : 1;
        const int per_pass = SEG ? (n_own < kBrickSeg / dl ? n_own : (kBrickSeg / dl > 1 ? kBrickSeg / dl : 1))
                                 : n_own;
        // the stencil of staged centre particle a: its halo cell hc, the x and
        // y offsets [lo, hi], and its z cells as at most two runs of
        // consecutive halo cells, whose particles are consecutive in the
        // staging order
        auto stencil = [&](int a, int (&hc)[3], int (&lo)[2], int (&hi)[2], int& za0, int& za1, int& zb0,
                           int& zb1) {
            split(s_hc[a], hc[0], hc[1], hc[2]);
            for (int q = 0; q < 2; ++q) {
                const bool w = (whole >> q) & 1;
                lo[q] = w ? (gd[q] >= 3 ? -1 : 0) : -1;
                hi[q] = w ? (gd[q] >= 2 ? 1 : 0) : 1;
            }
            zb0 = zb1 = 0;
            if (!((whole >> 2) & 1)) {
                za0 = hc[2] - 1;
                za1 = hc[2] + 2;
            } else if (gd[2] <= 3) {
                za0 = 0;
                za1 = gd[2];
            } else if (hc[2] == 0) {
                za0 = 0;
                za1 = 2;
                zb0 = gd[2] - 1;
                zb1 = gd[2];
            } else if (hc[2] == gd[2] - 1) {
                za0 = 0;
                za1 = 1;
                zb0 = gd[2] - 2;
                zb1 = gd[2];
            } else {
                za0 = hc[2] - 1;
                za1 = hc[2] + 2;
            }
        };
        for (int p0 = 0; p0 < n_own; p0 += per_pass) {
        const int p1 = n_own - p0 < per_pass ? n_own : p0 + per_pass;
        const int seg_cap = kBrickSeg / (p1 - p0);
        if (SEG && seg_cap >= d.list_cap) {
            // Dense bricks, every list in an LDS column: T lanes per centre
            // particle share its nine stencil rows (a brick has fewer centre
            // particles than the block has threads) and append the
            // FP32-accepted candidates to the particle's column through a
            // counter in s_cen (free once the staging is done).  Then the
            // FP64 test and the ordering by slot id, so the list does not
            // depend on the order of the appends: a wave per particle (lists
            // up to 128 entries), else one lane per particle sorting its
            // column.
            const int npp = p1 - p0;
            const int T = BS / npp < 1 ? 1 : (BS / npp > 4 ? 4 : BS / npp);
            int* const s_cnt = s_cen;
            for (int k = threadIdx.x; k < npp; k += BS) s_cnt[k] = 0;
            lds_barrier();
            for (int kk = threadIdx.x; kk < npp * T; kk += BS) {
                const int k = p0 + kk / T, sub = kk % T;
                const int a = s_cidx[k];
                int hc[3], lo[2], hi[2], za0, za1, zb0, zb1;
                stencil(a, hc, lo, hi, za0, za1, zb0, zb1);
                const int ti = MULTI ? (int)s_ty[a] : 0;
                const float4 pi = s_pf[a];
                const float xf = pi.x, yf = pi.y, zf = pi.z;
                int32_t* const col = L.seg + (k - p0);
                int* const cnt = s_cnt + (k - p0);
                const int cap = d.list_cap;
                // uw: std::true_type when the FP32 copies are unwrapped
                auto scan = [&](auto uw, int b0, int b1) {
                    constexpr bool UW = decltype(uw)::value;
                    // the next candidate's position is read while this one is
                    // tested (the index stays in the run)
                    float4 pn = s_pf[b0 < b1 ? b0 : 0];
                    for (int b2 = b0; b2 < b1; ++b2) {
                        // FP32 test first: it never rejects a pair the FP64 test keeps
                        const float4 pj = pn;
                        pn = s_pf[b2 + 1 < b1 ? b2 + 1 : b2];
                        const float ex = UW ? pj.x - xf : min_image_f(pj.x - xf, bxf);
                        const float ey = UW ? pj.y - yf : min_image_f(pj.y - yf, byf);
                        const float ez = UW ? pj.z - zf : min_image_f(pj.z - zf, bzf);
                        if (!((ex * ex + ey * ey) + ez * ez <= rl2f)) continue;
                        // a particle is staged once per region: b2 == a is j == i
                        if (b2 == a) continue;
                        if (MULTI && (int)s_ty[b2] < ti) continue;
                        // past the capacity a candidate is only counted (an
                        // overflow: the chunk is replayed with a larger one)
                        const int pos = atomicAdd(cnt, 1);
                        if (pos < cap) col[pos * npp] = b2;
                    }
                };
                auto rows = [&](auto uw) {
                    int row_i = 0;  // the particle's stencil rows dealt out over its T lanes
                    for (int o0 = lo[0]; o0 <= hi[0]; ++o0) {
                        const int g0 = ((whole & 1) ? (hc[0] + o0 + gd[0]) % gd[0] : hc[0] + o0);
                        for (int o1 = lo[1]; o1 <= hi[1]; ++o1) {
                            if (row_i++ % T != sub) continue;
                            const int g1 = ((whole & 2) ? (hc[1] + o1 + gd[1]) % gd[1] : hc[1] + o1);
                            const int row = (g0 * hd[1] + g1) * hd[2];
                            scan(uw, s_off[row + za0], s_off[row + za1]);
                            if (zb1 > zb0) scan(uw, s_off[row + zb0], s_off[row + zb1]);
                        }
                    }
                };
                if (unwrap)
                    rows(std::true_type{});
                else
                    rows(std::false_type{});
            }
            lds_barrier();
            if (d.dense && d.list_cap <= 128) {
                // one wave per centre particle: lane q runs the FP64 test of
                // column entries q and q + 64, and each kept slot id goes
                // straight to its sorted position, its rank among the kept
                // ids (distinct), counted over the wave's lanes (sparse
                // systems, whose columns hold a few entries, keep one lane
                // per particle: a wave per particle would serialize them)
                // PP particles per wave and iteration, their loads and
                // tests interleaved (each particle's chain of dependent LDS
                // reads would otherwise run alone).
                constexpr int NW = BS / 64, PP = 4;
                const int wv = threadIdx.x >> 6, wl = threadIdx.x & 63;
                const int cap = d.list_cap;
                for (int k0 = p0 + wv; k0 < p1; k0 += PP * NW) {
                    int n[PP], m[PP];
                    int32_t ii[PP];
                    double xi[PP], yi[PP], zi[PP];
                    const int32_t* colp[PP];
#pragma unroll
                    for (int p = 0; p < PP; ++p) {
                        const bool has = k0 + p * NW < p1;
                        const int k = has ? k0 + p * NW : k0;
                        const int a = s_cidx[k];
                        ii[p] = s_id[a];
                        xi[p] = s_x[a];
                        yi[p] = s_y[a];
                        zi[p] = s_z[a];
                        n[p] = has ? s_cnt[k - p0] : 0;
                        m[p] = n[p] < cap ? n[p] : cap;
                        colp[p] = L.seg + (k - p0);
                    }
                    auto kept_id = [&](int p, int q) -> uint32_t {  // ~0u: rejected
                        const int b2 = colp[p][q * npp];
                        const double dx = min_image_sel(s_x[b2] - xi[p], bx);
                        const double dy = min_image_sel(s_y[b2] - yi[p], by);
                        const double dz = min_image_sel(s_z[b2] - zi[p], bz);
                        const double r2 = (dx * dx + dy * dy) + dz * dz;
                        return r2 <= rl2 ? (uint32_t)s_id[b2] : ~0u;
                    };
                    uint32_t v0[PP], v1[PP];
#pragma unroll
                    for (int p = 0; p < PP; ++p) {
                        v0[p] = wl < m[p] ? kept_id(p, wl) : ~0u;
                        v1[p] = wl + 64 < m[p] ? kept_id(p, wl + 64) : ~0u;
                    }
                    // ranks of the kept ids (lanes past a particle's m hold
                    // ~0u, below no id)
                    int r0[PP], r1[PP];
                    int mx = 0;
#pragma unroll
                    for (int p = 0; p < PP; ++p) {
                        r0[p] = r1[p] = 0;
                        mx = m[p] > mx ? m[p] : mx;
                    }
                    if (mx <= 64) {
                        // one entry per lane: a bitonic sort over the wave
                        // (ids ascending, the ~0u of rejected entries last),
                        // after which lane r holds the entry of rank r
#pragma unroll
                        for (int kb = 2; kb <= 64; kb <<= 1) {
#pragma unroll
                            for (int jb = kb >> 1; jb > 0; jb >>= 1) {
                                const bool keep_min = ((wl & kb) == 0) == ((wl & jb) == 0);
#pragma unroll
                                for (int p = 0; p < PP; ++p) {
                                    const uint32_t o = (uint32_t)__shfl_xor((int)v0[p], jb, 64);
                                    v0[p] = keep_min ? (o < v0[p] ? o : v0[p]) : (o > v0[p] ? o : v0[p]);
                                }
                            }
                        }
#pragma unroll
                        for (int p = 0; p < PP; ++p) r0[p] = wl;
                    } else {
                        const int mx0 = mx < 64 ? mx : 64;
                        for (int l = 0; l < mx0; ++l) {
#pragma unroll
                            for (int p = 0; p < PP; ++p) {
                                const uint32_t x = __builtin_amdgcn_readlane(v0[p], l);
                                r0[p] += x < v0[p] ? 1 : 0;
                                r1[p] += x < v1[p] ? 1 : 0;
                            }
                        }
                    }
                    for (int l = 64; l < mx; ++l) {
#pragma unroll
                        for (int p = 0; p < PP; ++p) {
                            const uint32_t x = __builtin_amdgcn_readlane(v1[p], l - 64);
                            r0[p] += x < v0[p] ? 1 : 0;
                            r1[p] += x < v1[p] ? 1 : 0;
                        }
                    }
#pragma unroll
                    for (int p = 0; p < PP; ++p) {
                        if (k0 + p * NW >= p1) break;
                        int32_t* const col = d.nbr + ii[p];
                        if (v0[p] != ~0u) col[(int64_t)r0[p] * N] = (int32_t)v0[p];
                        if (v1[p] != ~0u) col[(int64_t)r1[p] * N] = (int32_t)v1[p];
                        const int kept = __popcll(__ballot(v0[p] != ~0u)) + __popcll(__ballot(v1[p] != ~0u));
                        // more FP32-accepted candidates than the capacity
                        // counts as an overflow (the chunk is replayed with
                        // a larger one)
                        if (wl == 0) list_close(d, ii[p], n[p] > cap ? n[p] : kept);
                    }
                }
            } else
            for (int k = p0 + threadIdx.x; k < p1; k += BS) {
                const int a = s_cidx[k];
                const int32_t i = s_id[a];
                const double xi = s_x[a], yi = s_y[a], zi = s_z[a];
                const int n = s_cnt[k - p0];
                RegList<SEG> lst(d.nbr + i, N, d.list_cap);
                lst.seg = L.seg + (k - p0);
                lst.seg_stride = npp;
                // the FP64 test (the reference's minimum image) over the
                // FP32-accepted candidates, in a loop every lane runs at once
                int kept = 0;
                const int m = n < d.list_cap ? n : d.list_cap;
                // the next entry is read before this one is kept (kept <= q)
                int bn = m > 0 ? lst.seg[0] : 0;
                for (int q = 0; q < m; ++q) {
                    const int b2 = bn;
                    if (q + 1 < m) bn = lst.seg[(q + 1) * lst.seg_stride];
                    const double dx = min_image_sel(s_x[b2] - xi, bx);
                    const double dy = min_image_sel(s_y[b2] - yi, by);
                    const double dz = min_image_sel(s_z[b2] - zi, bz);
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (r2 <= rl2) lst.seg[kept++ * lst.seg_stride] = s_id[b2];
                }
                lst.store_seg(kept);
                list_close(d, i, n > d.list_cap ? n : kept);
            }
        } else
        for (int k = p0 + threadIdx.x; k < p1; k += BS) {
            const int a = s_cidx[k];
            int hc[3], lo[2], hi[2], za0, za1, zb0, zb1;
            stencil(a, hc, lo, hi, za0, za1, zb0, zb1);
            const int32_t i = s_id[a];
            const int ti = MULTI ? (int)s_ty[a] : 0;
            const double xi = s_x[a], yi = s_y[a], zi = s_z[a];
            const float4 pi = s_pf[a];
            const float xf = pi.x, yf = pi.y, zf = pi.z;
            RegList<SEG> lst(d.nbr + i, N, d.list_cap);
            // uw: std::true_type when the FP32 copies are unwrapped
            auto scan = [&](auto uw, int b0, int b1) {
                constexpr bool UW = decltype(uw)::value;
                // the next candidate's position is read while this one is
                // tested (the index stays in the run)
                float4 pn = s_pf[b0 < b1 ? b0 : 0];
                for (int b2 = b0; b2 < b1; ++b2) {
                    // FP32 test first: it never rejects a pair the FP64 test keeps
                    const float4 pj = pn;
                    pn = s_pf[b2 + 1 < b1 ? b2 + 1 : b2];
                    const float ex = UW ? pj.x - xf : min_image_f(pj.x - xf, bxf);
                    const float ey = UW ? pj.y - yf : min_image_f(pj.y - yf, byf);
                    const float ez = UW ? pj.z - zf : min_image_f(pj.z - zf, bzf);
                    if (!((ex * ex + ey * ey) + ez * ez <= rl2f)) continue;
                    // a particle is staged once per region: b2 == a is j == i
                    if (b2 == a) continue;
                    if (MULTI && (int)s_ty[b2] < ti) continue;
                    const int32_t j = s_id[b2];
                    const double dx = min_image_sel(s_x[b2] - xi, bx);
                    const double dy = min_image_sel(s_y[b2] - yi, by);
                    const double dz = min_image_sel(s_z[b2] - zi, bz);
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (!(r2 <= rl2)) continue;
                    lst.push(j);
                }
            };
            auto rows = [&](auto uw) {
                for (int o0 = lo[0]; o0 <= hi[0]; ++o0) {
                    const int g0 = ((whole & 1) ? (hc[0] + o0 + gd[0]) % gd[0] : hc[0] + o0);
                    for (int o1 = lo[1]; o1 <= hi[1]; ++o1) {
                        const int g1 = ((whole & 2) ? (hc[1] + o1 + gd[1]) % gd[1] : hc[1] + o1);
                        const int row = (g0 * hd[1] + g1) * hd[2];
                        scan(uw, s_off[row + za0], s_off[row + za1]);
                        if (zb1 > zb0) scan(uw, s_off[row + zb0], s_off[row + zb1]);
                    }
                }
            };
            if (unwrap)
                rows(std::true_type{});
            else
                rows(std::false_type{});
            lst.store();
            list_close(d, i, lst.n);
        }
        if (p1 < n_own) lds_barrier();  // the next pass reuses the columns
        }  // passes
        }  // the region fitted in LDS
        lds_barrier();
        // the next region: the sub-brick's next cell, the brick's next
        // sub-brick, else the next brick
        if (level == 2) {
            if (++j2 < n2) continue;
            level = 1;
        }
        if (level == 1 && ++j1 < n1) continue;
        break;
        }  // regions
    }
}

// ---------------------------------------------------------------- fused step

// LDS ordering within one wave: a compiler fence (a wave's LDS operations
// execute in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// LDS a wave's list forces use (owner positions and slots, pair terms and
// flags, the owner map of a round).
template <typename FT>
constexpr int wave_scratch_bytes() {
    return 64 * (3 * (int)sizeof(double) + (int)sizeof(typename Acc<FT>::Terms) + 2 * (int)sizeof(int32_t) + 1);
}

// The list forces of a wave's particles (potential.rs:158-216 over the Verlet
// lists at the drifted positions).  In sparse systems (!WALK) each lane takes
// its particle's first entry, whose record the caller gathered a tile ahead
// (o0, type tj0); the remaining entries of the wave are spread over its 64
// lanes, one entry per lane and round, so the index and record loads of a
// round go out together instead of one entry at a time per particle.  In
// dense systems (WALK) the rounds cover every entry, or,
// with long lists, each lane walks its own list.  Each lane computes its
// pair's terms into the wave's LDS scratch, then every particle adds its own
// terms in list order: the order of a per-particle loop, so the sums are
// bitwise those of one.  A neighbour's drift is recomputed from its step
// record (position and kicked velocity) with the operations its own thread
// performs.
template <typename FT, bool MULTI, bool WALK>
__device__ __forceinline__ Acc<FT> wave_list_forces(const DevState& d, const CRec* __restrict__ src, char* wreg,
                                                    int n_mine, const CRec& o0, int tj0, int32_t e0, int64_t s,
                                                    double x1, double y1, double z1, int ti, bool split, double lam,
                                                    double dt, double bx, double by, double bz,
                                                    const PairParam& p0) {
    typedef typename Acc<FT>::Terms Terms;
    constexpr int W = 64;
    constexpr int K0 = WALK ? 0 : 1;  // first list position the rounds cover
    double* ox = (double*)wreg;
    double* oy = ox + W;
    double* oz = oy + W;
    Terms* pt = (Terms*)(oz + W);
    int32_t* os = (int32_t*)(pt + W);
    uint32_t* omap = (uint32_t*)(os + W);  // entry of a round -> owner lane << 16 | list position
    uint8_t* pf = (uint8_t*)(omap + W);
    const int wl = threadIdx.x & (W - 1);
    Acc<FT> acc;
    auto pair = [&](const CRec& o, int tj, double xi, double yi, double zi, Terms& tm, bool drift) -> bool {
        double xj = o.x, yj = o.y, zj = o.z;
        if (drift) {
            xj = wrap1(xj + (o.ux * lam) * dt, bx);
            yj = wrap1(yj + (o.uy * lam) * dt, by);
            zj = wrap1(zj + (o.uz * lam) * dt, bz);
        }
        const double dx = min_image(xj - xi, bx);
        const double dy = min_image(yj - yi, by);
        const double dz = min_image(zj - zi, bz);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        const PairParam& pp = MULTI ? d.pairs[ti * d.n_types + tj] : p0;
        if (!(r2 <= pp.r2max)) return false;
        tm = Acc<FT>::terms(dx, dy, dz, r2, pp);
        return true;
    };
    auto gathered = [&](int32_t j, double xi, double yi, double zi, Terms& tm) -> bool {
        if constexpr (WALK) {
            // dense systems: the neighbour's drifted position, written by
            // k_dpos for this launch with the operations its own thread uses
            const double4 q = d.dpos[j];
            CRec o;
            o.x = q.x;
            o.y = q.y;
            o.z = q.z;
            o.ux = o.uy = o.uz = 0.0;
            return pair(o, MULTI ? (int)d.slot_type[j] : 0, xi, yi, zi, tm, false);
        } else {
            const CRec o = src[j];  // one copy: the record loads as whole 16-byte words
            return pair(o, MULTI ? (int)d.slot_type[j] : 0, xi, yi, zi, tm, !split);
        }
    };
    if (!WALK && n_mine >= 1) {
        Terms tm;
        if (pair(o0, tj0, x1, y1, z1, tm, !split)) acc.add_terms(tm);
    }
    const int n_rest = n_mine > K0 ? n_mine - K0 : 0;
    if (!WALK && !__ballot(n_rest > 0)) return acc;
    const int incl = wave_scan_incl(n_rest);
    const int total = __builtin_amdgcn_readlane(incl, W - 1);
    if (WALK && (!__ballot(n_mine > 1) || total >= kLaneWalk * W)) {
        // at most one entry per lane, or long lists: each lane walks its own
        // list; entry k of the wave's particles is one coalesced row of the
        // column-major lists
        for (int k = 0; k < n_mine; ++k) {
            const int32_t j = k == 0 ? e0 : d.nbr[(int64_t)k * d.n_slots + s];
            Terms tm;
            if (gathered(j, x1, y1, z1, tm)) acc.add_terms(tm);
        }
        return acc;
    }
    const int my_off = incl - n_rest;
    wave_lds_sync();  // the previous tile's rounds are done with the scratch
    ox[wl] = x1;
    oy[wl] = y1;
    oz[wl] = z1;
    os[wl] = (int32_t)s;
    for (int c0 = 0; c0 < total; c0 += W) {
        // this round's entries [c0, c0 + W): the owner marks its own
        const int a0 = my_off > c0 ? my_off : c0;
        const int a1 = my_off + n_rest < c0 + W ? my_off + n_rest : c0 + W;
        for (int a = a0; a < a1; ++a) omap[a - c0] = ((uint32_t)wl << 16) | (uint32_t)(a - my_off + K0);
        wave_lds_sync();
        Terms tm{};
        uint8_t ok = 0;
        if (c0 + wl < total) {
            const uint32_t om = omap[wl];
            const int own = (int)(om >> 16), k = (int)(om & 0xffffu);
            // the index from the list rows (row 0 too): a choice between the
            // first entry in LDS and a row in global memory would compile to
            // one flat load, which waits on both memory counters
            const int32_t j = d.nbr[(int64_t)k * d.n_slots + os[own]];
            ok = gathered(j, ox[own], oy[own], oz[own], tm) ? 1 : 0;
        }
        pt[wl] = tm;
        pf[wl] = ok;
        wave_lds_sync();
        for (int a = a0; a < a1; ++a)
            if (pf[a - c0]) acc.add_terms(pt[a - c0]);
        wave_lds_sync();
    }
    return acc;
}

// One launch of the step kernel over the block's tiles, on the step records
// (position and kicked velocity).  The record, list length and first list
// entry of the next tile are loaded before the current tile is processed; in
// sparse systems the record of that first entry is gathered as the tile
// starts.
// The last step of a chunk also writes the full records (x, v, F) and the
// per-particle energies.
template <typename FT, int BS, bool MULTI, bool DENSE>
__global__ __launch_bounds__(BS, 4) void k_step(DevState d) {
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int gi = d.blk_group[b];
    const GroupInfo G = d.groups[gi];
    const int r = G.rep;
    const RepCtrl& C = ctrl_in(d, r);
    const uint32_t flags = C.flags;
    if (flags & (F_EXIT | F_BUILD)) {  // a build launch builds in the finalize kernel
        prof_block(d, P_STEP, 0, false, 0);
        return;
    }
    const uint64_t t_start = prof_now(d);
    const RepParams& P = d.params[r];
    const RepInfo R = d.reps[r];
    const TileRange tr = block_tiles(d, G, b, BS);
    const int lane = threadIdx.x;
    const int64_t base = G.start + lane;
    // step-record buffers (the full records' buffers are cur and full_dst)
    const int cur = (C.cur ^ C.cflip) & 1, nx = cur ^ 1;
    const CRec* __restrict__ rin = d.crec[cur];
    CRec* __restrict__ rnx = d.crec[nx];
    if (flags & F_NULL) {
        for (int64_t t = tr.t0; t < tr.t1; ++t) {
            const int64_t s = base + t * BS;
            if (t * BS + lane < G.count) grid_insert(d, C, R, s, rin[s].x, rin[s].y, rin[s].z);
        }
        __syncthreads();
        prof_block(d, P_STEP, t_start, true, 0);
        return;
    }
    const double bx = wave_uniform(C.box[0]), by = wave_uniform(C.box[1]), bz = wave_uniform(C.box[2]);
    const double dt = wave_uniform(P.dt), lam = wave_uniform(C.lambda), mu = wave_uniform(C.myu);
    const double m = wave_uniform(G.mass);
    const int ti = G.type;
    const double c = wave_uniform(P.cm[ti]);
    const double inv_m2 = wave_uniform(P.inv_m2[ti]);
    // the thermal sums are taken about the last stored centre-of-mass velocity
    const double* gs = d.sums + (int64_t)gi * NS;
    const double vr0 = wave_uniform(sums_vref(gs, G.sum_m, 0));
    const double vr1 = wave_uniform(sums_vref(gs, G.sum_m, 1));
    const double vr2 = wave_uniform(sums_vref(gs, G.sum_m, 2));
    const uint32_t kick_cols = P.th_kind == 2 ? kColsKick : 0u;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    if (flags & F_SPLIT_A) {
        // first half of a split step: kick, scale, drift, wrap into cur ^ 1
        // (the velocity field now holds the scaled velocity) and fill the grid
        // with the drifted positions (the drift exceeds the skin)
        for (int64_t t = tr.t0; t < tr.t1; ++t) {
            const int64_t s = base + t * BS;
            if (t * BS + lane >= G.count) continue;
            CRec a = rin[s];
            const double vkx = a.ux, vky = a.uy, vkz = a.uz;
            part[S_KVX] += vkx * m;
            part[S_KVY] += vky * m;
            part[S_KVZ] += vkz * m;
            const double wx = vkx - vr0, wy = vky - vr1, wz = vkz - vr2;
            part[S_KV2] += m * ((wx * wx + wy * wy) + wz * wz);
            a.ux = vkx * lam;
            a.uy = vky * lam;
            a.uz = vkz * lam;
            a.x = wrap1(a.x + a.ux * dt, bx);
            a.y = wrap1(a.y + a.uy * dt, by);
            a.z = wrap1(a.z + a.uz * dt, bz);
            rnx[s] = a;
            part[S_VMAX] = fmax(part[S_VMAX], (a.ux * a.ux + a.uy * a.uy) + a.uz * a.uz);
            grid_insert(d, C, R, s, a.x, a.y, a.z);
        }
        partial_store<BS>(d, part, b, kick_cols | kColsDrift);
        prof_block(d, P_STEP, t_start, true, 0);
        return;
    }
    // second half of a split step: drifted state in cur ^ 1, results to cur
    const bool split = (flags & F_SPLIT_B) != 0;
    const CRec* __restrict__ src = split ? rnx : rin;
    CRec* __restrict__ dst = split ? d.crec[cur] : rnx;
    // the full records of the chunk's last step go to the buffer not holding
    // the chunk's starting state
    Rec<FT>* __restrict__ full = (Rec<FT>*)d.rec[P.full_dst & 1];
    const PairParam p0 = wave_uniform(d.pairs[0]);
    int visits = 0;  // list entries read by this thread (profiling)
    // A wave's 64 records move between HBM and LDS as whole 16-byte words
    // with consecutive lanes on consecutive words (fully coalesced); each
    // lane then reads and writes its own record in the wave's part of the
    // LDS tile, so only the wave itself orders these accesses (no block
    // barrier).  The next tile's words load into registers while the current
    // tile is processed.
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    constexpr int RQ = (int)(sizeof(CRec) / 16);
    constexpr int WS = wave_scratch_bytes<FT>();
    __shared__ u4 s_tile[BS * RQ];
    __shared__ u4 s_ws[(BS / 64) * WS / 16];
    char* const wreg = (char*)s_ws + (threadIdx.x >> 6) * WS;
    const int wl = lane & 63, w64 = lane & ~63;  // lane in the wave, the wave's first lane
    u4* const wtile = s_tile + w64 * RQ;          // the wave's 64 records
    u4 pre[RQ];
    int nn = 0;
    // the step record (and type) of the first list entry of the thread's
    // particle, gathered a tile ahead (entry 0 of every slot exists)
    int32_t e0 = 0;
    if (tr.t0 < tr.t1) {
        const u4* g = (const u4*)(src + G.start + tr.t0 * BS + w64);
#pragma unroll
        for (int q = 0; q < RQ; ++q) pre[q] = g[q * 64 + wl];
        nn = tr.t0 * BS + lane < G.count ? d.ncount[base + tr.t0 * BS] : 0;
        e0 = d.nbr[base + tr.t0 * BS];
    }
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = base + t * BS;
        const bool valid = t * BS + lane < G.count;
        // sparse systems: the record of the particle's first list entry
        CRec o0;
        int tj0 = 0;
        if constexpr (!DENSE) {
            if (valid && nn > 0) {
                o0 = src[e0];
                if (MULTI) tj0 = d.slot_type[e0];
            } else {
                o0.x = o0.y = o0.z = o0.ux = o0.uy = o0.uz = 0.0;
            }
        }
#pragma unroll
        for (int q = 0; q < RQ; ++q) wtile[q * 64 + wl] = pre[q];
        wave_lds_sync();
        const CRec a = ((const CRec*)wtile)[wl];
        int nnn = 0;
        int32_t ne0 = 0;
        if (t + 1 < tr.t1) {
            const u4* g = (const u4*)(src + G.start + (t + 1) * BS + w64);
#pragma unroll
            for (int q = 0; q < RQ; ++q) pre[q] = g[q * 64 + wl];
            nnn = (t + 1) * BS + lane < G.count ? d.ncount[s + BS] : 0;
            ne0 = d.nbr[s + BS];
        }
        // a padding slot (past the group's count) keeps the padding record
        CRec res;
        res.x = res.y = res.z = __builtin_nan("");
        res.ux = res.uy = res.uz = 0.0;
        double x1 = 0.0, y1 = 0.0, z1 = 0.0, v1x = 0.0, v1y = 0.0, v1z = 0.0;
        if (valid) {
            if (split) {
                x1 = a.x;
                y1 = a.y;
                z1 = a.z;
                v1x = a.ux;
                v1y = a.uy;
                v1z = a.uz;
            } else {
                // integrator.rs:28-34 first half kick: the record holds it
                const double vkx = a.ux, vky = a.uy, vkz = a.uz;
                part[S_KVX] += vkx * m;
                part[S_KVY] += vky * m;
                part[S_KVZ] += vkz * m;
                const double wx = vkx - vr0, wy = vky - vr1, wz = vkz - vr2;
                part[S_KV2] += m * ((wx * wx + wy * wy) + wz * wz);
                // thermostat.rs:55-57 scale, integrator.rs:40-45 drift + wrap
                v1x = vkx * lam;
                v1y = vky * lam;
                v1z = vkz * lam;
                x1 = wrap1(a.x + v1x * dt, bx);
                y1 = wrap1(a.y + v1y * dt, by);
                z1 = wrap1(a.z + v1z * dt, bz);
                part[S_VMAX] = fmax(part[S_VMAX], (v1x * v1x + v1y * v1y) + v1z * v1z);
            }
        }
        const Acc<FT> acc = wave_list_forces<FT, MULTI, DENSE>(d, src, wreg, valid ? nn : 0, o0, tj0, e0, s, x1, y1,
                                                               z1, ti, split, lam, dt, bx, by, bz, p0);
        if (valid) {
            // integrator.rs:47-53 second half kick, barostat.rs:45-48 scale
            const double vx = v1x + (double)acc.fx * c;
            const double vy = v1y + (double)acc.fy * c;
            const double vz = v1z + (double)acc.fz * c;
            res.x = x1 * mu;
            res.y = y1 * mu;
            res.z = z1 * mu;
            // the next step's first half kick, as integrator.rs:28-34 rounds it
            res.ux = vx + (double)acc.fx * c;
            res.uy = vy + (double)acc.fy * c;
            res.uz = vz + (double)acc.fz * c;
            if (flags & F_WRITE_PV) {
                d.pot[s] = (double)acc.pot;
                d.vir[s] = (double)acc.vir;
                Rec<FT> o{};
                o.x = res.x;
                o.y = res.y;
                o.z = res.z;
                o.vx = vx;
                o.vy = vy;
                o.vz = vz;
                o.fx = acc.fx;
                o.fy = acc.fy;
                o.fz = acc.fz;
                o.type = (uint32_t)ti;
                full[s] = o;
            }
            part[S_MVX] += vx * m;
            part[S_MVY] += vy * m;
            part[S_MVZ] += vz * m;
            {
                const double wx = vx - vr0, wy = vy - vr1, wz = vz - vr2;
                part[S_MV2] += m * ((wx * wx + wy * wy) + wz * wz);
            }
            part[S_POT] += (double)acc.pot;
            part[S_VIR] += (double)acc.vir;
            part[S_WV] = fmax(part[S_WV], (vx * vx + vy * vy) + vz * vz);
            part[S_WA] = fmax(part[S_WA],
                              (((double)acc.fx * (double)acc.fx + (double)acc.fy * (double)acc.fy) +
                               (double)acc.fz * (double)acc.fz) * inv_m2);
            if (flags & F_INSERT) grid_insert(d, C, R, s, res.x, res.y, res.z);
        } else if (flags & F_WRITE_PV) {
            Rec<FT> o{};
            o.x = o.y = o.z = __builtin_nan("");
            o.type = (uint32_t)ti;
            full[s] = o;
        }
        visits += valid ? nn : 0;
        wave_lds_sync();  // every lane has read its input record
        ((CRec*)wtile)[wl] = res;
        wave_lds_sync();
        {
            u4* g = (u4*)(dst + G.start + t * BS + w64);
#pragma unroll
            for (int q = 0; q < RQ; ++q) g[q * 64 + wl] = wtile[q * 64 + wl];
        }
        wave_lds_sync();  // the tile is read out before the next one overwrites it
        nn = nnn;
        e0 = ne0;
    }
    __shared__ int s_vis[BS / 64];
    if (d.prof) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) visits += __shfl_xor(visits, o, 64);
        if ((lane & 63) == 0) s_vis[lane >> 6] = visits;
    }
    // a split step stored its kick / drift columns in the first half
    partial_store<BS>(d, part, b, split ? kColsState : (kColsState | kick_cols | kColsDrift));
    if (d.prof && lane == 0) {
        uint64_t tot = 0;
        for (int w = 0; w < BS / 64; ++w) tot += (uint64_t)s_vis[w];
        prof_block(d, P_STEP, t_start, true, tot);
    }
}

// Step records of the list path from the full records of each replica's
// current buffer: the position and the first half kick of the next step
// (integrator.rs:28-34).
template <typename FT, int BS>
__global__ __launch_bounds__(BS) void k_pack(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const double c = d.params[G.rep].cm[G.type];
    const RepCtrl& C = ctrl_in(d, G.rep);
    const TileRange tr = block_tiles(d, G, b, BS);
    const Rec<FT>* __restrict__ in = (const Rec<FT>*)d.rec[C.cur];
    CRec* __restrict__ out = d.crec[(C.cur ^ C.cflip) & 1];
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + threadIdx.x;
        const Rec<FT> a = in[s];
        CRec o;
        o.x = a.x;
        o.y = a.y;
        o.z = a.z;
        o.ux = a.vx + (double)a.fx * c;
        o.uy = a.vy + (double)a.fy * c;
        o.uz = a.vz + (double)a.fz * c;
        out[s] = o;
    }
}

// Stand-alone list build (md_update_force on the list paths): replicas whose
// control block is flagged F_BUILD build their lists from grid `fill_epoch`.
template <typename FT, bool MULTI>
__global__ __launch_bounds__(kFinBS) void k_build(DevState d) {
    const int b = blockIdx.x;
    const int r = d.bld_rep[b];
    const RepCtrl& C = ctrl_in(d, r);
    if ((C.flags & (F_BUILD | F_EXIT)) != F_BUILD) return;
    const RepInfo R = d.reps[r];
    brick_build<Rec<FT>, kFinBS, MULTI, true>(d, C, d.params[r], R, (const Rec<FT>*)d.rec[C.build_src],
                                        b - R.bld_first, R.bld_count);
}

// Force evaluation at the current positions through the current lists
// (update_force for the list paths).
template <typename FT, int BS, bool MULTI>
__global__ __launch_bounds__(BS) void k_force_eval(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const int r = G.rep;
    const RepCtrl& C = ctrl_in(d, r);
    const TileRange tr = block_tiles(d, G, b, BS);
    Rec<FT>* __restrict__ rc = (Rec<FT>*)d.rec[C.cur];
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double m = G.mass;
    const int ti = G.type;
    const double* gs = d.sums + (int64_t)d.blk_group[b] * NS;
    const double vr[3] = {sums_vref(gs, G.sum_m, 0), sums_vref(gs, G.sum_m, 1), sums_vref(gs, G.sum_m, 2)};
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + threadIdx.x;
        if (t * BS + threadIdx.x >= G.count) continue;
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
        part[S_MVX] += a.vx * m;
        part[S_MVY] += a.vy * m;
        part[S_MVZ] += a.vz * m;
        {
            const double wx = a.vx - vr[0], wy = a.vy - vr[1], wz = a.vz - vr[2];
            part[S_MV2] += m * ((wx * wx + wy * wy) + wz * wz);
        }
        part[S_POT] += (double)acc.pot;
        part[S_VIR] += (double)acc.vir;
        part[S_WV] = fmax(part[S_WV], (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz);
        part[S_WA] = fmax(part[S_WA], (((double)acc.fx * (double)acc.fx + (double)acc.fy * (double)acc.fy) +
                                       (double)acc.fz * (double)acc.fz) / (m * m));
    }
    partial_store<BS>(d, part, b, kColsAll);
}

template <typename FT, int BS>
__global__ __launch_bounds__(BS) void k_insert_all(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = ctrl_in(d, G.rep);
    const RepInfo R = d.reps[G.rep];
    const TileRange tr = block_tiles(d, G, b, BS);
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + threadIdx.x;
        if (t * BS + threadIdx.x >= G.count) continue;
        const Rec<FT>& a = ((const Rec<FT>*)d.rec[C.cur])[s];
        grid_insert(d, C, R, s, a.x, a.y, a.z);
    }
}

// Partial sums of the current state (after an upload or a host-side change).
template <typename FT, int BS>
__global__ __launch_bounds__(BS) void k_reduce_state(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = ctrl_in(d, G.rep);
    const TileRange tr = block_tiles(d, G, b, BS);
    const double* gs = d.sums + (int64_t)d.blk_group[b] * NS;
    const double vr[3] = {sums_vref(gs, G.sum_m, 0), sums_vref(gs, G.sum_m, 1), sums_vref(gs, G.sum_m, 2)};
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + threadIdx.x;
        if (t * BS + threadIdx.x >= G.count) continue;
        const Rec<FT> a = ((const Rec<FT>*)d.rec[C.cur])[s];
        const double m = G.mass;
        part[S_MVX] += a.vx * m;
        part[S_MVY] += a.vy * m;
        part[S_MVZ] += a.vz * m;
        const double wx = a.vx - vr[0], wy = a.vy - vr[1], wz = a.vz - vr[2];
        part[S_MV2] += m * ((wx * wx + wy * wy) + wz * wz);
        part[S_POT] += d.pot[s];
        part[S_VIR] += d.vir[s];
        const double fx = (double)a.fx, fy = (double)a.fy, fz = (double)a.fz;
        part[S_WV] = fmax(part[S_WV], (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz);
        part[S_WA] = fmax(part[S_WA], ((fx * fx + fy * fy) + fz * fz) / (m * m));
    }
    partial_store<BS>(d, part, b, kColsAll);
}

// ---------------------------------------------------------------- O(N^2) FP64 path

// integrator.rs:28-45: kick, scale, drift, wrap into buffer cur ^ 1.
template <int BS>
__global__ __launch_bounds__(BS) void k_drift(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const RepCtrl& C = ctrl_in(d, G.rep);
    if (C.flags & F_EXIT) {
        prof_block(d, P_DRIFT, 0, false, 0);
        return;
    }
    const uint64_t t_start = prof_now(d);
    const RepParams& P = d.params[G.rep];
    const TileRange tr = block_tiles(d, G, b, BS);
    const int cur = C.cur, nx = cur ^ 1;
    const double dt = P.dt, lam = C.lambda, m = G.mass;
    const double c = P.cm[G.type];
    const double* gs = d.sums + (int64_t)d.blk_group[b] * NS;
    const double vr[3] = {sums_vref(gs, G.sum_m, 0), sums_vref(gs, G.sum_m, 1), sums_vref(gs, G.sum_m, 2)};
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + threadIdx.x;
        if (t * BS + threadIdx.x >= G.count) continue;
        Rec<double> a = ((const Rec<double>*)d.rec[cur])[s];
        const double vkx = a.vx + a.fx * c;
        const double vky = a.vy + a.fy * c;
        const double vkz = a.vz + a.fz * c;
        part[S_KVX] += vkx * m;
        part[S_KVY] += vky * m;
        part[S_KVZ] += vkz * m;
        const double wx = vkx - vr[0], wy = vky - vr[1], wz = vkz - vr[2];
        part[S_KV2] += m * ((wx * wx + wy * wy) + wz * wz);
        a.vx = vkx * lam;
        a.vy = vky * lam;
        a.vz = vkz * lam;
        a.x = wrap1(a.x + a.vx * dt, C.box[0]);
        a.y = wrap1(a.y + a.vy * dt, C.box[1]);
        a.z = wrap1(a.z + a.vz * dt, C.box[2]);
        ((Rec<double>*)d.rec[nx])[s] = a;
        part[S_VMAX] = fmax(part[S_VMAX], (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz);
    }
    partial_store<BS>(d, part, b, (P.th_kind == 2 ? kColsKick : 0u) | kColsDrift);
    prof_block(d, P_DRIFT, t_start, true, 0);
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
    const RepCtrl& C = ctrl_in(d, r);
    if (!FORCE_ONLY && (C.flags & F_EXIT)) {
        prof_block(d, P_STEP, 0, false, 0);
        return;
    }
    const uint64_t t_start = prof_now(d);
    const RepParams& P = d.params[r];
    const RepInfo R = d.reps[r];
    const TileRange tr = block_tiles(d, G, b, BS);
    const int cur = C.cur;
    const Rec<double>* __restrict__ src = (const Rec<double>*)d.rec[FORCE_ONLY ? cur : (cur ^ 1)];
    Rec<double>* __restrict__ dst = (Rec<double>*)d.rec[cur];
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const int ti = G.type;
    const int T = d.n_types;
    const double m = G.mass;
    const double c = P.cm[ti];
    const double* gs = d.sums + (int64_t)d.blk_group[b] * NS;
    const double vr[3] = {sums_vref(gs, G.sum_m, 0), sums_vref(gs, G.sum_m, 1), sums_vref(gs, G.sum_m, 2)};
    __shared__ double tx[BS], ty[BS], tz[BS];
    __shared__ uint8_t tt[BS];
    const PairParam p0 = d.pairs[ti * T + ti];
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    for (int64_t it = tr.t0; it < tr.t1; ++it) {
        const int64_t s = G.start + it * BS + threadIdx.x;
        const bool valid = it * BS + threadIdx.x < G.count;
        Rec<double> a;
        if (valid) a = src[s];
        double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0, vir = 0.0;
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
        if (valid) {
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
            part[S_MVX] += a.vx * m;
            part[S_MVY] += a.vy * m;
            part[S_MVZ] += a.vz * m;
            const double wx = a.vx - vr[0], wy = a.vy - vr[1], wz = a.vz - vr[2];
            part[S_MV2] += m * ((wx * wx + wy * wy) + wz * wz);
            part[S_POT] += pot;
            part[S_VIR] += vir;
            part[S_WV] = fmax(part[S_WV], (a.vx * a.vx + a.vy * a.vy) + a.vz * a.vz);
            part[S_WA] = fmax(part[S_WA], ((fx * fx + fy * fy) + fz * fz) / (m * m));
        }
    }
    partial_store<BS>(d, part, b, FORCE_ONLY ? kColsAll : kColsState);
    if (!FORCE_ONLY) prof_block(d, P_STEP, t_start, true, 0);
}

// ---------------------------------------------------------------- finalize

__device__ __forceinline__ void thermo(const double* S, const GroupInfo& G, const double* box,
                                       bool kicked, double& temperature, double& pressure) {
    // sum m |v - v_com|^2 (energy.rs:25-37, pressure.rs:11-14): the reduced
    // sums hold it already, taken about the centre of mass
    const double th2 = S[kicked ? S_KV2 : S_MV2];
    const double thermal = th2 / 2.0;
    // temperature.rs:4-7
    temperature = (2.0 * thermal) / (3.0 * (double)G.count * kKB) * 100.0;
    // pressure.rs:9-19
    const double volume = box[0] * box[1] * box[2];
    pressure = (th2 + (-S[S_VIR]) * 0.5) / volume / 3.0;
}

__device__ __forceinline__ void fill_geometry(double (&gbox)[3], int32_t (&gdim)[3], int32_t (&bedge)[3],
                                              int32_t& bwhole, const RepParams& P, const RepInfo& R,
                                              const double* gb) {
#pragma unroll
    for (int k = 0; k < 3; ++k) gbox[k] = gb[k];
    grid_dims(gbox, P.rlist, R.cell_budget, gdim);
    brick_dims(gdim, R.n_atoms, bedge, &bwhole);
}

// Geometry of the next grid fill (buffer ins_epoch & 1), chosen with fixed
// indices so that the control block can live in registers.
__device__ __forceinline__ void set_fill(RepCtrl& C, const RepParams& P, const RepInfo& R, const double* gb) {
    C.ins_epoch = C.fill_epoch + 1u;
    if (C.ins_epoch & 1)
        fill_geometry(C.gbox[1], C.gdim[1], C.bedge[1], C.bwhole[1], P, R, gb);
    else
        fill_geometry(C.gbox[0], C.gdim[0], C.bedge[0], C.bwhole[0], P, R, gb);
}

// Advance a replica's controls after a launch (M_STEP) or prepare the first
// step of a call (M_PRIME).  S holds the group sums of the replica's types.
__device__ __forceinline__ void advance(RepCtrl& C, const RepParams& P, const RepInfo& R, const GroupInfo* G,
                        const double* S, int mode) {
    const int T = R.n_types;
    const double dt = P.dt;
    const bool baro = P.ba_kind == 1;
    // the list terms that depend on the sums alone, computed first so that
    // their square roots overlap the control chain below (the same
    // operations as where they are used: the values are bitwise the same)
    double vmax = 0.0, wmax = 0.0;
    for (int t = 0; t < T; ++t) {
        vmax = fmax(vmax, S[t * NS + S_VMAX]);
        wmax = fmax(wmax, (sqrt(S[t * NS + S_WV]) + (dt / 2.0) * sqrt(S[t * NS + S_WA])) * (1.0 + 1e-12));
    }
    const double vmax_root = sqrt(vmax);
    // lists built in this launch describe the current positions exactly
    const bool built = mode == M_STEP && (C.flags & F_BUILD);
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
            C.build_src = (C.cur ^ C.cflip) ^ 1;
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
                    C.D += vmax_root * (1.0 + 1e-12) * dt / C.s;
                }
                if (baro) C.s *= C.myu;
            }
            if (f & F_INSERT) {
                C.fill_epoch = C.ins_epoch;
                nf |= F_GRID_FRESH;
            }
            // k_step wrote buffer cur ^ 1 (a split step, and every surrogate
            // step, ends in cur)
            if (P.list_mode && !P.in_place && !(f & F_SPLIT_B)) C.cur ^= 1;
            // the chunk's last step wrote the full records to buffer full_dst:
            // the next chunk's step records are packed from there
            if (P.full_dst >= 0 && C.step == P.target_step) {
                C.cflip = (C.cur ^ C.cflip) ^ P.full_dst;  // the step records stay where they are
                C.cur = P.full_dst;
            }
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
        // (wmax above)
        const double w = C.lambda * wmax * dt / C.s;
        const double thr = P.rlist - P.rc_max / C.s - P.margin;
        const bool need = !(C.flags & F_HAVE_LIST) || (P.rebuild_every_step && !built) ||
                          !(2.0 * (C.D + w) < thr);
        if (need) {
            const double w0 = C.lambda * wmax * dt;
            if (!(2.0 * w0 < P.rlist - P.rc_max - P.margin)) {
                // one drift exceeds what any list built before it can cover
                C.flags |= F_SPLIT_A;
                set_fill(C, P, R, C.box);
            } else if (C.flags & F_GRID_FRESH) {
                C.flags |= F_BUILD;
                C.build_src = C.cur ^ C.cflip;
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
__device__ __forceinline__ void reduce_rows(const DevState& d, const GroupInfo& G, int lo, int hi, double* out) {
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

// Word-wise block copy (LDS staging of the control structures).
template <int BS>
__device__ __forceinline__ void copy_words(void* dst, const void* src, int bytes) {
    for (int i = threadIdx.x; i < bytes / 4; i += BS) ((uint32_t*)dst)[i] = ((const uint32_t*)src)[i];
}

// Finalize of replica r by one block: group sums (from the block partials of
// this launch, or the stored ones when the launch did not step), then one
// thread advances the controls on register copies of the LDS-staged inputs,
// and the block writes the control block of the next launch.  Returns the
// launch index when the replica did work in this launch, else -1.
__device__ __forceinline__ int64_t finalize_replica(const DevState& d, int r, int mode) {
    constexpr int BS = kFinBS;
    const int T = d.n_types;
    const int64_t g0 = (int64_t)r * T;  // groups are replica-major
    __shared__ RepCtrl s_ctrl;
    __shared__ RepParams s_par;
    __shared__ GroupInfo s_grp[kMaxTypes];
    __shared__ double s_sums[kMaxTypes * NS];
    __shared__ double s_vref[kMaxTypes * 3];
    const RepInfo R = d.reps[r];
    // one system of one type: the column totals come from the exact
    // accumulators; their loads go out together with the staging loads
    // (one memory round trip)
    const bool single = d.acc_mode != 0;
    double acc_k = 0.0;
    if (single && mode != M_PRIME && (int)threadIdx.x < NS) acc_k = acc_total(d, threadIdx.x);
    copy_words<BS>(&s_ctrl, &ctrl_in(d, r), (int)sizeof(RepCtrl));
    copy_words<BS>(&s_par, d.params + r, (int)sizeof(RepParams));
    copy_words<BS>(s_grp, d.groups + g0, (int)sizeof(GroupInfo) * T);
    // the reference velocities this launch's thermal partials were taken
    // about: the centre of mass of the sums stored so far
    if ((int)threadIdx.x < 3 * T) {
        const int t = threadIdx.x / 3;
        s_vref[threadIdx.x] = sums_vref(d.sums + (g0 + t) * NS, d.groups[g0 + t].sum_m, threadIdx.x % 3);
    }
    __syncthreads();
    const uint32_t f0 = s_ctrl.flags;
    const int64_t launch = s_ctrl.launches;
    if (mode == M_STEP && (f0 & F_EXIT)) {
        // target reached: carry the block over to the buffer the next launch reads
        if (threadIdx.x == 0) s_ctrl.launches++;
        __syncthreads();
        copy_words<BS>(ctrl_out(d, r), &s_ctrl, (int)sizeof(RepCtrl));
        return -1;
    }
    // the new group sums are stored at the end: a store before a barrier would
    // make every thread wait for its completion
    const bool reduced = finalize_reduces(mode, f0);
    if (reduced) {
        if (single) {
            if ((int)threadIdx.x < NS) {
                s_sums[threadIdx.x] = acc_k;
                acc_clear(d, threadIdx.x);  // for the next launch's partials
            }
        } else {
            for (int t = 0; t < T; ++t) reduce_rows<BS>(d, s_grp[t], 0, s_grp[t].nblocks, s_sums + t * NS);
        }
    } else {
        for (int i = threadIdx.x; i < T * NS; i += BS) s_sums[i] = d.sums[g0 * NS + i];
    }
    __syncthreads();
    if (reduced && (int)threadIdx.x < T) {
        // thermal sums moved from the reference velocity to the centre of mass
        double* S = s_sums + threadIdx.x * NS;
        const double M = s_grp[threadIdx.x].sum_m;
        const double* vr = s_vref + 3 * threadIdx.x;
        S[S_MV2] = sums_about_com(S[S_MV2], S + S_MVX, M, vr);
        S[S_KV2] = sums_about_com(S[S_KV2], S + S_KVX, M, vr);
    }
    __syncthreads();
    if (d.mlp_stat) {
        // surrogate: the LJ force check of this step, and fallbacks
        MlpStat* ms = d.mlp_stat + r;
        const bool stepped = mode == M_STEP && finalize_reduces(mode, f0);
        const int je = d.mlp_c->judge_every;
        if (stepped && je > 0 && (s_ctrl.step + 1) % je == 0) {
            __shared__ double s_j[NS];
            double jv[NS];
#pragma unroll
            for (int k = 0; k < NS; ++k) jv[k] = 0.0;
            for (int q = threadIdx.x; q < R.nblocks; q += BS) {
                jv[0] += d.mlp_judge[R.first_block + q];
                jv[1] += d.mlp_judge[d.n_blocks + R.first_block + q];
            }
            block_store<BS>(jv, s_j, 1, 3u);
            if (threadIdx.x == 0) {
                ms->last_err2 = s_j[0];
                ms->last_ref2 = s_j[1];
                if (s_j[1] > 0.0) ms->max_rel = fmax(ms->max_rel, sqrt(s_j[0] / s_j[1]));
                ms->checks++;
            }
        }
        if (threadIdx.x == 0 && ms->oor) {
            if (stepped) ms->fallbacks++;
            ms->oor = 0u;
        }
    }
    if (mode == M_REDUCE) {
        for (int i = threadIdx.x; i < T * NS; i += BS) d.sums[g0 * NS + i] = s_sums[i];
        return -1;
    }
    if (threadIdx.x == 0) {
        RepCtrl C = s_ctrl;
        const RepParams P = s_par;
        advance(C, P, R, s_grp, s_sums, mode);
        s_ctrl = C;
    }
    __syncthreads();
    copy_words<BS>(ctrl_out(d, r), &s_ctrl, (int)sizeof(RepCtrl));
    if (reduced)
        for (int i = threadIdx.x; i < T * NS; i += BS) d.sums[g0 * NS + i] = s_sums[i];
    return mode == M_STEP ? launch : -1;
}

// Fold the per-block records of this launch's per-slot kernels into the
// ring entry of the launch (one block).
__device__ __forceinline__ void prof_fold(const DevState& d, int64_t launch, bool drift) {
    constexpr int BS = kFinBS;
    __shared__ unsigned long long s_mn[BS / 64], s_mx[BS / 64], s_en[BS / 64];
    for (int slot = 0; slot < (drift ? 2 : 1); ++slot) {
        const int k = slot == 0 ? P_STEP : P_DRIFT;
        const unsigned long long* e = d.prof_blk + (int64_t)k * d.n_blocks * 3;
        unsigned long long mn = ~0ull, mx = 0ull, en = 0ull;
        for (int i = threadIdx.x; i < d.n_blocks; i += BS) {
            mn = e[3 * i] < mn ? e[3 * i] : mn;
            mx = e[3 * i + 1] > mx ? e[3 * i + 1] : mx;
            en += e[3 * i + 2];
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const unsigned long long a = __shfl_xor(mn, o, 64), b = __shfl_xor(mx, o, 64);
            mn = a < mn ? a : mn;
            mx = b > mx ? b : mx;
            en += __shfl_xor(en, o, 64);
        }
        if ((threadIdx.x & 63) == 0) {
            s_mn[threadIdx.x >> 6] = mn;
            s_mx[threadIdx.x >> 6] = mx;
            s_en[threadIdx.x >> 6] = en;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < BS / 64; ++w) {
                mn = s_mn[w] < mn ? s_mn[w] : mn;
                mx = s_mx[w] > mx ? s_mx[w] : mx;
                en += s_en[w];
            }
            if (mn != ~0ull) {
                unsigned long long* r = d.prof + ((launch % kProfRing) * P_NUM + k) * 2;
                r[0] = mn;
                r[1] = mx;
            }
            d.prof[kProfCount] += en;
        }
        __syncthreads();
    }
    if (d.mlp_w && threadIdx.x == 0) {
        for (int k : {(int)P_NET, (int)P_FEAT}) {
            unsigned long long* e = d.prof_blk + (int64_t)k * d.n_blocks * 3;
            if (e[0] != ~0ull) {
                unsigned long long* r = d.prof + ((launch % kProfRing) * P_NUM + k) * 2;
                r[0] = e[0];
                r[1] = e[1];
            }
            e[0] = ~0ull;
            e[1] = 0ull;
        }
    }
}

template <typename FT, bool MULTI, bool BUILD, bool REPS, bool DENSE>
__global__ __launch_bounds__(kFinBS) void k_fin(DevState d, int mode) {
    const int b = blockIdx.x;
    const uint64_t t_start = prof_now(d);
    // the last block folds the per-block records of k_step (or k_drift +
    // k_allpairs) while the first blocks finalize
    if (d.prof && mode == M_STEP && b == (int)gridDim.x - 1)
        prof_fold(d, ctrl_in(d, 0).launches, !BUILD || d.mlp_w != nullptr);
    int64_t launch = -1;
    if (b < d.n_reps) launch = finalize_replica(d, b, mode);
    if (BUILD && !REPS && b < d.n_bld) {
        // one system: the build blocks share its bricks
        const RepCtrl& C = ctrl_in(d, 0);
        if ((C.flags & (F_BUILD | F_EXIT)) == F_BUILD) {
            const RepInfo R = d.reps[0];
            if (d.compact)
                brick_build<CRec, kFinBS, MULTI, DENSE>(d, C, d.params[0], R, d.crec[C.build_src], b, d.n_bld);
            else
                brick_build<Rec<FT>, kFinBS, MULTI, DENSE>(d, C, d.params[0], R, (const Rec<FT>*)d.rec[C.build_src], b,
                                                    d.n_bld);
            launch = C.launches;
        }
    } else if (BUILD && b < d.n_bld) {
        // several replicas: the bricks of every replica building in this
        // launch, numbered replica after replica, are shared out among all
        // build blocks: block b takes bricks b, b + n_bld, ... (replicas
        // rebuild at different launches, so a share fixed per replica would
        // leave most blocks idle while a large replica builds).  A block's
        // bricks within one replica are one strided run: one builder call
        // per (replica, block).
        __shared__ int s_rb[kFinBS + 1];
        __shared__ int s_work[kFinBS];
        __shared__ int s_nwork;
        const int nb = d.n_bld;
        int64_t base = 0;  // bricks of the replicas before this chunk
        for (int r0 = 0; r0 < d.n_reps; r0 += kFinBS) {
            const int n = d.n_reps - r0 < kFinBS ? d.n_reps - r0 : kFinBS;
            if ((int)threadIdx.x < n) {
                const RepCtrl& C = ctrl_in(d, r0 + threadIdx.x);
                s_rb[threadIdx.x] = (C.flags & (F_BUILD | F_EXIT)) == F_BUILD ? brick_count(C) : 0;
            }
            if (threadIdx.x == 0) s_nwork = 0;
            lds_barrier();
            block_exclusive_scan<kFinBS>(s_rb, n);
            const int tot = __builtin_amdgcn_readfirstlane(s_rb[n]);  // uniform: scalar registers
            // replicas holding bricks of this block (in any order: each
            // brick's lists depend on the brick alone)
            if ((int)threadIdx.x < n) {
                const int64_t lo = base + s_rb[threadIdx.x], hi = base + s_rb[threadIdx.x + 1];
                const int64_t g0 = lo + ((b - lo % nb) % nb + nb) % nb;
                if (g0 < hi) s_work[atomicAdd(&s_nwork, 1)] = threadIdx.x;
            }
            lds_barrier();
            const int nwork = __builtin_amdgcn_readfirstlane(s_nwork);
            for (int w = 0; w < nwork; ++w) {
                const int idx = __builtin_amdgcn_readfirstlane(s_work[w]);
                const int64_t lo = base + __builtin_amdgcn_readfirstlane(s_rb[idx]);
                const int r = r0 + idx;
                const int bi0 = (int)(((b - lo % nb) % nb + nb) % nb);
                const RepCtrl& C = ctrl_in(d, r);
                const RepInfo R = d.reps[r];
                if (d.compact)
                    brick_build<CRec, kFinBS, MULTI, DENSE>(d, C, d.params[r], R, d.crec[C.build_src], bi0, nb);
                else
                    brick_build<Rec<FT>, kFinBS, MULTI, DENSE>(d, C, d.params[r], R,
                                                               (const Rec<FT>*)d.rec[C.build_src], bi0, nb);
                launch = C.launches;
            }
            base += tot;
            lds_barrier();  // s_rb and s_work are read before the next chunk rewrites them
        }
    }
    if (launch >= 0) {
        __syncthreads();
        prof_mark(d, launch, P_FIN, t_start);
    }
}

// ---------------------------------------------------------------- small helpers

template <typename FT>
__global__ __launch_bounds__(256) void k_scale(DevState d, int64_t start, int64_t count,
                                               double fac, int which) {
    const int64_t i = start + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= start + count) return;
    Rec<FT>& a = ((Rec<FT>*)d.rec[ctrl_in(d, 0).cur])[i];
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
    const RepCtrl& C = ctrl_in(d, G.rep);
    const TileRange tr = block_tiles(d, G, b, BS);
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + threadIdx.x;
        if (t * BS + threadIdx.x >= G.count) continue;
        Rec<FT>& a = ((Rec<FT>*)d.rec[C.cur])[s];
        a.x = wrap1(a.x, C.box[0]);
        a.y = wrap1(a.y, C.box[1]);
        a.z = wrap1(a.z, C.box[2]);
    }
}

// Timestamp ring reset: starts at the largest tick, ends at zero; the list
// entry counter after the ring at zero.
__global__ __launch_bounds__(256) void k_prof_init(DevState d) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < kProfCount) d.prof[i] = (i & 1) ? 0ull : ~0ull;
    if (i == kProfCount) d.prof[i] = 0ull;
    // the atomic {min start, max end} pairs of k_mlp_net and the feature phase
    if (i == kProfCount + 1 && d.prof_blk) d.prof_blk[(int64_t)P_NET * d.n_blocks * 3] = ~0ull;
    if (i == kProfCount + 2 && d.prof_blk) d.prof_blk[(int64_t)P_NET * d.n_blocks * 3 + 1] = 0ull;
    if (i == kProfCount + 3 && d.prof_blk) d.prof_blk[(int64_t)P_FEAT * d.n_blocks * 3] = ~0ull;
    if (i == kProfCount + 4 && d.prof_blk) d.prof_blk[(int64_t)P_FEAT * d.n_blocks * 3 + 1] = 0ull;
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

#define MD_LAUNCH_FT_MULTI(K, B)                                                               \
    if (fp32) {                                                                                \
        if (multi) hipLaunchKernelGGL((K<float, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((K<float, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    } else {                                                                                   \
        if (multi) hipLaunchKernelGGL((K<double, B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
        else hipLaunchKernelGGL((K<double, B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);      \
    }

// Drifted positions of a dense launch: the step's drift of every slot
// (integrator.rs:40-45 with the thermostat scale), computed as k_step
// computes its own particle's, or the record's position in the second half
// of a split step.  Padding slots keep their NaN positions.
template <int BS>
__global__ __launch_bounds__(BS) void k_dpos(DevState d) {
    const int b = blockIdx.x;
    const int gi = d.blk_group[b];
    const GroupInfo G = d.groups[gi];
    const RepCtrl& C = ctrl_in(d, G.rep);
    const uint32_t flags = C.flags;
    if (flags & (F_EXIT | F_BUILD | F_NULL | F_SPLIT_A)) return;
    const RepParams& P = d.params[G.rep];
    const TileRange tr = block_tiles(d, G, b, BS);
    const int cur = (C.cur ^ C.cflip) & 1;
    const bool split = (flags & F_SPLIT_B) != 0;
    const CRec* __restrict__ src = d.crec[split ? cur ^ 1 : cur];
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double dt = P.dt, lam = C.lambda;
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + threadIdx.x;
        const CRec a = src[s];
        double4 q;
        if (split || t * BS + (int)threadIdx.x >= G.count) {
            q.x = a.x;
            q.y = a.y;
            q.z = a.z;
        } else {
            q.x = wrap1(a.x + (a.ux * lam) * dt, bx);
            q.y = wrap1(a.y + (a.uy * lam) * dt, by);
            q.z = wrap1(a.z + (a.uz * lam) * dt, bz);
        }
        q.w = 0.0;
        d.dpos[s] = q;
    }
}

hipError_t launch_dpos(const DevState& d, int bs, hipStream_t st) {
#define L(B) hipLaunchKernelGGL((k_dpos<B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_step(const DevState& d, bool fp32, int bs, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define KS(FT, B, M)                                                                        \
    do {                                                                                    \
        if (d.dense)                                                                        \
            hipLaunchKernelGGL((k_step<FT, B, M, true>), dim3(d.n_blocks), dim3(B), 0, st, d);  \
        else                                                                                \
            hipLaunchKernelGGL((k_step<FT, B, M, false>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    } while (0)
#define L(B)                           \
    if (fp32) {                        \
        if (multi) KS(float, B, true);  \
        else KS(float, B, false);       \
    } else {                           \
        if (multi) KS(double, B, true); \
        else KS(double, B, false);      \
    }
    MD_BS_SWITCH(bs, L)
#undef L
#undef KS
    return hipGetLastError();
}

hipError_t launch_pack(const DevState& d, bool fp32, int bs, hipStream_t st) {
#define L(B)                                                                         \
    if (fp32)                                                                        \
        hipLaunchKernelGGL((k_pack<float, B>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else                                                                             \
        hipLaunchKernelGGL((k_pack<double, B>), dim3(d.n_blocks), dim3(B), 0, st, d);
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_build(const DevState& d, bool fp32, hipStream_t st) {
    const bool multi = d.n_types > 1;
    if (d.n_bld <= 0) return hipSuccess;
    if (fp32) {
        if (multi) hipLaunchKernelGGL((k_build<float, true>), dim3(d.n_bld), dim3(kFinBS), 0, st, d);
        else hipLaunchKernelGGL((k_build<float, false>), dim3(d.n_bld), dim3(kFinBS), 0, st, d);
    } else {
        if (multi) hipLaunchKernelGGL((k_build<double, true>), dim3(d.n_bld), dim3(kFinBS), 0, st, d);
        else hipLaunchKernelGGL((k_build<double, false>), dim3(d.n_bld), dim3(kFinBS), 0, st, d);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_prologue(DevState d, RepParams p, double* snap_sums, int n_groups,
                                                  MlpStat* snap_mlp, RepCtrl* snap_ctrl, int par) {
    const int t = threadIdx.x;
    // word copies (every structure is a whole number of 32-bit words)
    auto copy = [&](void* dst, const void* src, int bytes) {
        for (int i = t; i < bytes / 4; i += 256) ((uint32_t*)dst)[i] = ((const uint32_t*)src)[i];
    };
    copy((void*)d.params, &p, (int)sizeof(RepParams));
    for (int i = t; i < (int)(sizeof(RepStats) / 4); i += 256) ((uint32_t*)d.stats)[i] = 0u;
    if (!snap_ctrl) return;  // no snapshot in this prologue
    copy(snap_sums, d.sums, (int)sizeof(double) * NS * n_groups);
    if (snap_mlp) copy(snap_mlp, d.mlp_stat, (int)sizeof(MlpStat));
    copy(snap_ctrl, d.ctrl + (size_t)par * d.n_reps, (int)sizeof(RepCtrl));
}

__global__ __launch_bounds__(256) void k_epilogue(DevState d, int par, RepCtrl* h_ctrl, RepStats* h_stats) {
    const int t = threadIdx.x;
    const uint32_t* c = (const uint32_t*)(d.ctrl + (size_t)par * d.n_reps);
    for (int i = t; i < (int)(sizeof(RepCtrl) / 4) * d.n_reps; i += 256) ((uint32_t*)h_ctrl)[i] = c[i];
    for (int i = t; i < (int)(sizeof(RepStats) / 4); i += 256) ((uint32_t*)h_stats)[i] = ((const uint32_t*)d.stats)[i];
}

hipError_t launch_prologue(const DevState& d, const RepParams& p, double* snap_sums, int n_groups,
                           MlpStat* snap_mlp, RepCtrl* snap_ctrl, int par, hipStream_t st) {
    hipLaunchKernelGGL(k_prologue, dim3(1), dim3(256), 0, st, d, p, snap_sums, n_groups, snap_mlp, snap_ctrl, par);
    return hipGetLastError();
}

hipError_t launch_epilogue(const DevState& d, int par, RepCtrl* h_ctrl, RepStats* h_stats, hipStream_t st) {
    hipLaunchKernelGGL(k_epilogue, dim3(1), dim3(256), 0, st, d, par, h_ctrl, h_stats);
    return hipGetLastError();
}

hipError_t launch_prof_init(const DevState& d, hipStream_t st) {
    const int n = (int)kProfCount + 5;
    hipLaunchKernelGGL(k_prof_init, dim3((n + 255) / 256), dim3(256), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_force_eval(const DevState& d, bool fp32, int bs, hipStream_t st) {
    const bool multi = d.n_types > 1;
#define L(B) MD_LAUNCH_FT_MULTI(k_force_eval, B)
    MD_BS_SWITCH(bs, L)
#undef L
    return hipGetLastError();
}

hipError_t launch_finalize(const DevState& d, int mode, bool build, bool fp32, hipStream_t st) {
    const bool multi = d.n_types > 1;
    if (build && mode == M_STEP && d.n_bld > 0) {
        const int grid = d.n_bld > d.n_reps ? d.n_bld : d.n_reps;
        const bool reps = d.n_reps > 1;
#define KF2(FT, M, DN)                                                                                \
    do {                                                                                               \
        if (reps)                                                                                      \
            hipLaunchKernelGGL((k_fin<FT, M, true, true, DN>), dim3(grid), dim3(kFinBS), 0, st, d, mode);  \
        else                                                                                           \
            hipLaunchKernelGGL((k_fin<FT, M, true, false, DN>), dim3(grid), dim3(kFinBS), 0, st, d, mode); \
    } while (0)
#define KF(FT, M)                 \
    do {                          \
        if (d.dense)              \
            KF2(FT, M, true);     \
        else                      \
            KF2(FT, M, false);    \
    } while (0)
        if (fp32) {
            if (multi) KF(float, true);
            else KF(float, false);
        } else {
            if (multi) KF(double, true);
            else KF(double, false);
        }
#undef KF
#undef KF2
    } else {
        hipLaunchKernelGGL((k_fin<double, false, false, false, false>), dim3(d.n_reps), dim3(kFinBS), 0, st, d,
                           mode);
    }
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
