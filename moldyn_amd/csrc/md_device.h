// md_device.h -- device data layout and kernel launch interface of the
// MI355X step engine.  Shared by md_kernels.hip (device code) and md_api.hip
// (host runtime behind the C ABI).
//
// Layout in HBM (one context, R replicas x T types):
//  * particles live in "slots"; each (replica, type) group is a whole number
//    of tiles (one tile = one slot per thread of a block), and a block owns
//    `tpb` consecutive tiles of one group, so every block belongs to exactly
//    one group and the per-group FP64 reductions need no segmenting.
//    Padding slots hold NaN positions, which fail every distance test.
//  * state is an array of per-particle records (position and velocity FP64,
//    force FP32 on the fast path or FP64 on the exact paths), double-buffered:
//    a step reads buffer `cur` and writes `cur ^ 1`.  A neighbour's whole
//    record is one 64-byte gather (80 bytes with FP64 forces).
//  * Verlet lists are column-major (nbr[k * n_slots + slot]) so that the k-th
//    neighbour of 64 consecutive slots is one coalesced load.
//  * the cell grid is two epoch-tagged header arrays (count | epoch << 32)
//    with fixed-capacity slot arrays, so filling it never needs a clear and a
//    fill can run while the other buffer is being read.  Lists are built from
//    it brick by brick: a block stages a brick of cells plus a one-cell halo
//    (ids and positions) in LDS and builds the lists of the brick's atoms.
//  * the per-replica control blocks are double-buffered by launch parity:
//    every kernel of a launch reads buffer `par`, the finalize writes
//    `par ^ 1`, so blocks of the same kernel never see a half-updated block.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

namespace md {

constexpr int kMaxTypes = 8;
constexpr double kKB = 1.380648528;  // core/src/lib.rs:15

// per-particle state record
template <typename FT>
struct Rec;
template <>
struct alignas(16) Rec<float> {
    double x, y, z, vx, vy, vz;
    float fx, fy, fz;
    uint32_t type;
};
template <>
struct alignas(16) Rec<double> {
    double x, y, z, vx, vy, vz;
    double fx, fy, fz;
    uint32_t type, pad;
};
static_assert(sizeof(Rec<float>) == 64, "record size");
static_assert(sizeof(Rec<double>) == 80, "record size");

// The list path's step record inside md_step: the position and the kicked
// velocity u = v + F dt/(2m), the first half kick of the next step
// (integrator.rs:28-34), rounded as that kick rounds it.  The next step and
// every neighbour's drift need u alone, not v and F.  In the first half of a
// split step the velocity field holds the scaled velocity instead.  The full
// records (Rec) stay the state between calls: a chunk packs them at its
// start, and its last step writes them back.
struct alignas(16) CRec {
    double x, y, z, ux, uy, uz;
};
static_assert(sizeof(CRec) == 48, "step record size");

// per-launch control flags of a replica
enum : uint32_t {
    F_INSERT = 1u << 1,     // insert this launch's output positions into grid `ins_epoch`
    F_NULL = 1u << 2,       // no time step: insert the input positions
    F_EXIT = 1u << 3,       // target step reached: launch does nothing
    F_WRITE_PV = 1u << 4,   // store per-atom potential and virial in this launch
    F_HAVE_LIST = 1u << 5,  // lists exist (built at some fill)
    F_GRID_FRESH = 1u << 6, // grid `fill_epoch` holds the current input positions
    F_BAD = 1u << 7,        // non-finite T/P or mu^3 <= 0 seen
    F_SPLIT_A = 1u << 8,    // drift exceeds the skin: kick + drift + grid fill only
    F_SPLIT_B = 1u << 9,    // ... then force and kick at the drifted positions
    F_BUILD = 1u << 10,     // no time step: build the lists from grid `fill_epoch`
    F_SPLIT_PENDING = 1u << 11,  // a BUILD between the halves of a split step
};

// per-block partial / per-group sum columns
enum {
    S_MVX, S_MVY, S_MVZ,  // sum m v (state after the step)
    S_MV2,                // sum m |v - v_ref|^2 in partials; stored group sums: about v_com
    S_POT,                // sum potential
    S_VIR,                // sum virial (Particle::temp)
    S_KVX, S_KVY, S_KVZ,  // sum m v after the first half kick (Nose-Hoover)
    S_KV2,                // as S_MV2, after the first half kick
    S_WV,                 // max |v|^2 of the state after the step
    S_WA,                 // max |F|^2 / m^2 of the state after the step
    S_VMAX,               // max |v|^2 of this step's drift
    NS
};
constexpr int kFirstMax = S_WV;

struct PairParam {
    double sigma, eps, r_cut, u_cut;
    double r2max;  // largest double d with sqrt(d) <= r_cut
    float sig2f, eps24f, eps4f, ucutf;
};

struct GroupInfo {
    int64_t start, count;
    int32_t first_block, nblocks;
    int32_t rep, type;
    double mass;   // Particle::mass of the type
    double sum_m;  // sequential fold of the masses
};

struct RepInfo {
    int64_t start, end;  // slot range
    int32_t first_group, n_types;
    int32_t first_block, nblocks;
    int32_t bld_first, bld_count;  // k_build blocks of this replica
    int64_t cell_base;   // first cell of this replica in each grid buffer
    int64_t cell_budget; // cells available
    int64_t n_atoms;
};

struct RepParams {
    double dt;
    int32_t th_kind, ba_kind;  // thermostat 0/1 Berendsen/2 Nose-Hoover, barostat 0/1
    double th_tau, th_target, ba_beta, ba_tau, ba_target;
    int64_t target_step;
    int32_t list_mode;         // 1 when a Verlet list is used
    int32_t rebuild_every_step;
    double rlist, rc_max, margin, insert_growth;
    double psi_in;        // Nose-Hoover psi handed in by the caller
    int32_t use_psi_in;   // M_PRIME starts from psi_in instead of psi_used
    int32_t in_place;     // a step ends in the buffer it started from (surrogate path)
    // step-record path: the buffer the chunk's last step writes the full
    // records to, the one not holding the chunk's starting state (which a
    // replay after an overflow restarts from); -1 on the other paths
    int32_t full_dst;
    int32_t pad_p;
    double cm[kMaxTypes];     // dt / (2 m_t), the half-kick factor (integrator.rs:29)
    double inv_m2[kMaxTypes]; // 1 / m_t^2 (drift-bound bookkeeping only)
};

struct RepCtrl {
    double box[3];             // box at the start of the launch
    double lambda, myu, psi;   // controls of the next real step (psi after calculate_lambda)
    double lambda_used, myu_used, psi_used;  // of the last completed step (caller-visible)
    double s, D;               // scale since list build, drift bound in build units
    double gbox[2][3];         // box of grid buffer 0/1
    int32_t gdim[2][3];        // cells per axis of grid buffer 0/1
    int32_t bedge[2][3];       // brick edge in cells (= gdim on whole-axis bricks)
    int32_t bwhole[2];         // axes whose brick spans the whole grid (bit mask)
    uint32_t flags;
    uint32_t ins_epoch;        // epoch this launch inserts with
    uint32_t fill_epoch;       // epoch of the last completed fill
    int32_t cur;               // state buffer holding the input
    int32_t build_src;         // state buffer a BUILD launch reads (step records: their index)
    int32_t cflip;             // step records of the input: buffer cur ^ cflip
    int64_t step;              // completed real steps
    int64_t rebuilds, nulls, launches, splits;
};

struct RepStats {
    unsigned int list_overflow, cell_overflow;
    int max_list, max_cell;
    unsigned int brick_fallback, pad;
};

// Learned per-atom energy surrogate: 32 radial features, 128-128 SiLU MLP.
constexpr int kMlpK = 32;
constexpr int kMlpH = 128;
// flat parameter vector: W1 [H][K] | b1 [H] | W2 [H][H] | b2 [H] | w3 [H] | b3
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpH * kMlpK;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpH;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpH * kMlpH;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpH;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpH;
constexpr int kMlpNP = kMlpOffB3 + 1;
// BF16 inference copies: W1 [H][K] | W1t [K][H] | W2 [H][H] | W2t [H][H]
constexpr int kMlpBfW1 = 0;
constexpr int kMlpBfW1t = kMlpBfW1 + kMlpH * kMlpK;
constexpr int kMlpBfW2 = kMlpBfW1t + kMlpH * kMlpK;
constexpr int kMlpBfW2t = kMlpBfW2 + kMlpH * kMlpH;
constexpr int kMlpBfN = kMlpBfW2t + kMlpH * kMlpH;

struct MlpConst {
    float g_mean[kMlpK], g_inv_std[kMlpK];  // input normalisation
    float g_lo[kMlpK], g_hi[kMlpK];         // feature range seen in training
    float mu[kMlpK];                        // basis centres (nm)
    float inv_2w2;                          // a = 1 / (2 w^2)
    float inv_w2;                           // 1 / w^2
    // the centres are evenly spaced (mu_{k+1} = mu_k + D), so along k each
    // basis value is the previous one times q_k = exp(2 a D t_k - a D^2),
    // t_k = r - mu_k, and q_{k+1} = q_k exp(-2 a D^2)
    float bas_2ad, bas_ad2, bas_c;          // 2 a D, a D^2, exp(-2 a D^2)
    float rc;                               // basis cutoff (nm)
    float pi_rc;                            // pi / rc (cosine cutoff argument per nm)
    float e_scale, e_shift;
    float b3;
    int32_t judge_every;                    // LJ force check interval (steps), 0 = never
    int32_t check_range;                    // fall back to LJ outside [g_lo, g_hi]
};

// per-replica surrogate statistics (device)
struct MlpStat {
    double last_err2, last_ref2;  // sum |F - F_LJ|^2 and sum |F_LJ|^2 at the last check
    double max_rel;               // largest RMSE / RMS over the checks
    int64_t checks, fallbacks;
    uint32_t oor;                 // a feature left the training range in this launch
    uint32_t pad;
};

struct DevState {
    void* rec[2];            // Rec<float> or Rec<double>, n_slots each
    CRec* crec[2];           // step records of the list path (null on the other paths)
    double4* dpos;           // dense list path: every slot's drifted position (x, y, z, unused)
    const uint8_t* slot_type;  // type of every slot (padding slots: the group's type)
    int32_t compact;         // the list path steps on crec (k_fin builds from crec)
    int32_t dense;           // long lists expected: lanes walk their own lists, builds sort long lists in LDS
    double* pot;
    double* vir;
    int32_t* nbr;
    int32_t* ncount;
    int32_t list_cap;
    unsigned long long* cell_hdr[2];
    // slot k of cell c at cell_slots[buf][k * cell_plane + c] (k < cell_cap):
    // a plane per slot rank, so that the first slots of a run of cells are
    // contiguous (a sparse grid holds about one particle per cell, and a
    // brick reads every slot plane of its halo in runs of cells)
    int32_t* cell_slots[2];
    int32_t cell_cap;
    int64_t cell_plane;
    const int32_t* blk_group;
    const int32_t* bld_rep;  // k_build block -> replica
    const GroupInfo* groups;
    const RepInfo* reps;
    RepCtrl* ctrl;
    const RepParams* params;
    RepStats* stats;
    const PairParam* pairs;  // n_types x n_types
    double* partial;         // NS x n_blocks (column-major)
    // one system of one type: the per-block partials go instead to exact
    // fixed-point accumulators (kAccSeg segments x NS columns x 3 digits),
    // which the finalize reads and clears
    int32_t acc_mode;
    unsigned long long* acc;
    double* sums;            // n_groups x NS
    unsigned long long* prof;  // per-launch kernel timestamps (null: off)
    unsigned long long* prof_blk;  // per-block {start, end, list entries} of the per-slot kernels
    // surrogate path (null unless force mode MD_FORCE_MLP)
    const MlpConst* mlp_c;
    const float* mlp_theta;       // kMlpNP
    const __bf16* mlp_w;          // kMlpBfN
    __bf16* mlp_x;                // n_slots x kMlpK normalised features
    float* mlp_gam;               // n_slots x kMlpK  dE_i / dG_i
    __bf16* mlp_gamb;             // the same in BF16: the copy neighbours gather
    float4* mlp_pos;              // n_slots  FP32 drifted positions of the stepping launch
    int32_t* mlp_nbr;             // list_cap x n_slots  the list entries within r_c (column-major, as nbr)
    int32_t* mlp_ncount;          // n_slots  their number
    float* mlp_e;                 // n_slots          E_i
    MlpStat* mlp_stat;            // n_reps
    double* mlp_judge;            // 2 x n_blocks per-block check sums
    int32_t n_types;
    int32_t n_blocks;
    int32_t n_reps;
    int32_t par;             // control buffer this launch reads
    int32_t tpb;             // tiles per block of the per-slot kernels
    int32_t n_bld;           // list-build blocks
    int64_t n_slots;
};

__host__ __device__ inline const RepCtrl& ctrl_in(const DevState& d, int r) {
    return d.ctrl[d.par * d.n_reps + r];
}
__host__ __device__ inline RepCtrl* ctrl_out(const DevState& d, int r) {
    return d.ctrl + (d.par ^ 1) * d.n_reps + r;
}

// Tiles of block b: group G's tiles [t0, t1), slots G.start + t * BS + lane.
struct TileRange {
    int64_t t0, t1;
};
__host__ __device__ inline TileRange block_tiles(const DevState& d, const GroupInfo& G, int b, int bs) {
    const int64_t nt = (G.count + bs - 1) / bs;
    const int64_t t0 = (int64_t)(b - G.first_block) * d.tpb;
    const int64_t t1 = t0 + d.tpb < nt ? t0 + d.tpb : nt;
    return TileRange{t0, t1};
}

// Per-launch timestamp ring (wall clock ticks): for launch l, entry
// (l % kProfRing) holds {min start, max end} of kernels P_STEP, P_FIN,
// P_DRIFT, P_NET; one counter of list entries read follows the ring.  The
// per-slot kernels store per-block {start, end, entries} with plain stores
// (slot P_STEP for k_step / k_allpairs / k_mlp_force, P_DRIFT for k_drift /
// k_mlp_desc) and one block of the finalize kernel folds them into the ring;
// k_mlp_net (grid not one block per slot range) and the feature phase of
// k_mlp_desc (P_FEAT) keep {min start, max end} with atomics in the first
// two words of their per-block areas, which the fold moves to the ring and
// resets; the finalize kernel's own blocks (one per
// replica plus the blocks that build lists) record with atomics.
constexpr int kProfRing = 8192;
enum { P_STEP = 0, P_FIN = 1, P_DRIFT = 2, P_NET = 3, P_FEAT = 4, P_NUM = 5 };
constexpr int64_t kProfCount = (int64_t)kProfRing * P_NUM * 2;  // index of the entry counter

enum FinalizeMode { M_STEP = 0, M_REDUCE = 1, M_PRIME = 2 };

// the finalize stages these through LDS word by word
static_assert(sizeof(RepCtrl) % 4 == 0, "RepCtrl staging");
static_assert(sizeof(RepParams) % 4 == 0, "RepParams staging");
static_assert(sizeof(GroupInfo) % 4 == 0, "GroupInfo staging");

// threads per block of the finalize / list-build kernel
constexpr int kFinBS = 512;

// LDS capacity of one brick: halo cells and staged particles.  Dilute
// systems get bricks of up to 14^3 cells so that a brick still holds a few
// hundred particles.
constexpr int kBrickEdge = 14;
constexpr int kBrickCells = 4096;  // (14 + 2)^3
constexpr int kBrickAtoms = 1024;
constexpr int kBrickTarget = 900;  // expected staged particles a brick is sized for
constexpr int kBrickOwn = 440;     // expected centre particles: one per thread of a 512-thread block
constexpr int kListReg = 16;       // list entries kept (sorted) in registers while building
constexpr int kBrickSeg = 16384;   // LDS entries for lists longer than kListReg (per brick)
constexpr int kLaneWalk = 6;       // mean list entries per particle from which lanes walk their own lists

// Cell grid of a fill: per axis floor(L / (r_list (1 + 1e-9))) cells in
// [1, 1024], coarsened until the replica's cell budget holds.  Cell edges stay
// >= r_list, so the 27-cell stencil around a particle covers its list sphere.
// (Written with fixed indices only, so that the device control update keeps
// its state in registers.)
__host__ __device__ inline int coarsen(int g) {
    const int dec = g / 8;
    const int n = g - (dec < 1 ? 1 : dec);
    return n < 1 ? 1 : n;
}
__host__ __device__ inline void grid_dims(const double* gb, double rlist, int64_t budget, int* gd) {
    int g[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double q = floor(gb[k] / (rlist * (1.0 + 1e-9)));
        g[k] = q < 1.0 ? 1 : (q > 1024.0 ? 1024 : (int)q);
    }
    while ((int64_t)g[0] * g[1] * g[2] > budget) {
        // the largest axis, the first of equals
        if (g[1] > g[0] && g[1] >= g[2]) {
            g[1] = coarsen(g[1]);
        } else if (g[2] > g[0] && g[2] > g[1]) {
            g[2] = coarsen(g[2]);
        } else {
            g[0] = coarsen(g[0]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) gd[k] = g[k];
}

// Brick shape of a grid.  Candidates for e = 1..kBrickEdge are three shape
// families: (e, e, e), (e, e, e+1) and (e, e+1, e+1) cells.  An axis whose
// grid is no longer than the candidate edge + 2 is taken whole (no halo on
// it).  A candidate qualifies when its halo (edge + 2 cells per tiled axis)
// has at most kBrickCells cells and, at the grid's mean occupancy, at most
// kBrickTarget particles, and its centre at most kBrickOwn particles.  Of the
// qualifying shapes the one with the most centre cells wins, ties going to
// the smaller halo; with none, bricks are single cells.
__host__ __device__ inline void brick_dims(const int* gd, int64_t n_atoms, int* edge, int* whole) {
    const double per_cell = (double)n_atoms / ((double)gd[0] * gd[1] * gd[2]);
    const int g0 = gd[0], g1 = gd[1], g2 = gd[2];
    // edges e (cubic), then e with the last one or two axes a cell longer:
    // in dense systems a cubic brick that fits the staging capacity leaves
    // most threads without a centre particle, and one cell more along an
    // axis or two doubles the centre (scalars only: the device control
    // update keeps its state in registers)
    int b0 = 1, b1 = 1, b2 = 1;
    double best_own = -1.0, best_halo = 0.0;
    for (int e = 1; e <= kBrickEdge; ++e) {
        for (int x = 0; x < 3; ++x) {
            const int e0 = e, e1 = e + (x >= 2 ? 1 : 0), e2 = e + (x >= 1 ? 1 : 0);
            const double halo = (double)(g0 <= e0 + 2 ? g0 : e0 + 2) * (double)(g1 <= e1 + 2 ? g1 : e1 + 2) *
                                (double)(g2 <= e2 + 2 ? g2 : e2 + 2);
            const double own = (double)(g0 <= e0 + 2 ? g0 : e0) * (double)(g1 <= e1 + 2 ? g1 : e1) *
                               (double)(g2 <= e2 + 2 ? g2 : e2);
            if (halo <= (double)kBrickCells && halo * per_cell <= (double)kBrickTarget &&
                own * per_cell <= (double)kBrickOwn && (own > best_own || (own == best_own && halo < best_halo))) {
                best_own = own;
                best_halo = halo;
                b0 = e0;
                b1 = e1;
                b2 = e2;
            }
        }
    }
    int w = 0;
    edge[0] = g0 <= b0 + 2 ? g0 : b0;
    edge[1] = g1 <= b1 + 2 ? g1 : b1;
    edge[2] = g2 <= b2 + 2 ? g2 : b2;
    w |= g0 <= b0 + 2 ? 1 : 0;
    w |= g1 <= b1 + 2 ? 2 : 0;
    w |= g2 <= b2 + 2 ? 4 : 0;
    *whole = w;
}

// Exact accumulation of block partials (acc_mode): a sum column's value v is
// taken in units of 2^-60 (rounded to nearest there; |v| < 2^60) as a signed
// 120-bit integer split into three 40-bit digits, and each digit is added
// with an integer atomic into the accumulator of the block's segment.
// Integer sums do not depend on the order of the additions, so the result is
// the same whatever order the blocks finish in and however the particles are
// grouped into blocks.  Max columns take an atomic max of the bit pattern
// (non-negative doubles order like their bits).  A non-finite or too large
// value marks its column, whose sum is then NaN.
constexpr int kAccSeg = 16;
constexpr int kAccDigits = 3;
constexpr int kAccWords = kAccSeg * NS * kAccDigits + NS;  // + one mark word per column

// column masks for block partial stores
constexpr uint32_t kColsKick = (1u << S_KVX) | (1u << S_KVY) | (1u << S_KVZ) | (1u << S_KV2);
constexpr uint32_t kColsState = (1u << S_MVX) | (1u << S_MVY) | (1u << S_MVZ) | (1u << S_MV2) |
                                (1u << S_POT) | (1u << S_VIR) | (1u << S_WV) | (1u << S_WA);
constexpr uint32_t kColsDrift = 1u << S_VMAX;
constexpr uint32_t kColsAll = (1u << NS) - 1u;

// host-side launchers (md_kernels.hip)
hipError_t launch_step(const DevState& d, bool fp32, int bs, hipStream_t st);
hipError_t launch_pack(const DevState& d, bool fp32, int bs, hipStream_t st);
// Dense list path: the drifted positions of every slot for this launch's
// neighbour gathers (the step kernel then reads 32 bytes per list entry and
// recomputes no drift).
hipError_t launch_dpos(const DevState& d, int bs, hipStream_t st);
hipError_t launch_build(const DevState& d, bool fp32, hipStream_t st);
hipError_t launch_prof_init(const DevState& d, hipStream_t st);
// Chunk prologue of a one-replica context in one launch: the parameters
// (from the kernel argument), cleared statistics, and the snapshot a replay
// restarts from (group sums of n_groups groups, the surrogate's counters when
// snap_mlp is set, and the control block the next launch reads).
hipError_t launch_prologue(const DevState& d, const RepParams& p, double* snap_sums, int n_groups,
                           MlpStat* snap_mlp, RepCtrl* snap_ctrl, int par, hipStream_t st);
// Chunk epilogue: the control blocks the next launch reads (n_reps) and the
// statistics, written into host memory the device can address.
hipError_t launch_epilogue(const DevState& d, int par, RepCtrl* h_ctrl, RepStats* h_stats, hipStream_t st);
// finalize (grid n_reps); with `build` the grid also covers the list-build
// blocks, which build the lists of replicas flagged F_BUILD
hipError_t launch_finalize(const DevState& d, int mode, bool build, bool fp32, hipStream_t st);
hipError_t launch_drift(const DevState& d, int bs, hipStream_t st);
hipError_t launch_allpairs(const DevState& d, int bs, bool force_only, hipStream_t st);
hipError_t launch_insert_all(const DevState& d, bool fp32, int bs, hipStream_t st);
hipError_t launch_force_eval(const DevState& d, bool fp32, int bs, hipStream_t st);
hipError_t launch_reduce_state(const DevState& d, bool fp32, int bs, hipStream_t st);
hipError_t launch_scale(const DevState& d, bool fp32, int64_t start, int64_t count, double fac,
                        int which /*0 velocity, 1 position*/, hipStream_t st);
hipError_t launch_wrap(const DevState& d, bool fp32, int bs, hipStream_t st);

// surrogate step kernels (md_mlp.hip); force_only evaluates at the current
// positions without advancing time
hipError_t launch_mlp_desc(const DevState& d, int bs, bool force_only, int phase, hipStream_t st);
hipError_t launch_mlp_net(const DevState& d, int max_blocks, hipStream_t st);
hipError_t launch_mlp_force(const DevState& d, int bs, bool force_only, hipStream_t st);
hipError_t launch_mlp_convert(const DevState& d, hipStream_t st);

// initialisers on the device (md_init.hip)
hipError_t launch_init_blank(const DevState& d, bool fp32, const GroupInfo* groups, int n_groups, int64_t n_slots,
                             hipStream_t st);
hipError_t launch_init_lattice(void* rec, bool fp32, int64_t slot0, int64_t sites, bool fcc, const double start[3],
                               const int64_t grid[3], double a, hipStream_t st);
hipError_t launch_init_mb(void* rec, bool fp32, int64_t slot0, int64_t n, double sigma, uint64_t seed,
                          hipStream_t st);

}  // namespace md
