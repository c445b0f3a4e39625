// md_api.hip -- host runtime behind include/moldyn_hip.h.
//
// A context owns the slot layout, device buffers, one HIP stream and two
// captured hipGraphs of `graph_chunk` step launches (one per parity of the
// double-buffered control blocks).  md_step runs in chunks: snapshot the
// state (list paths), prime the controls on the device, replay the graph,
// let replicas that spent launches on list refreshes catch up, then check
// the device's overflow flags; on overflow the chunk is replayed from the
// snapshot with larger capacities (results do not depend on when lists are
// rebuilt, because every list is sorted and pairs outside r_cut contribute
// nothing).  Capacities that come close to overflowing are grown between
// chunks without a replay.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "md_device.h"
#include "md_frames.h"
#include "moldyn_hip.h"

using namespace md;

static thread_local std::string g_error;

namespace {

enum { K_STEP = 0, K_FINALIZE = 1, K_ALLPAIRS = 2, K_DRIFT = 3, K_BUILD = 4, K_MLP_NET = 5, K_MLP_FEAT = 6, K_NUM = 7 };

}  // namespace

struct md_ctx {
    int device = 0;
    hipStream_t st = nullptr;
    int R = 1, T = 1;
    std::vector<int64_t> counts;  // R x T
    std::vector<double> masses;   // T
    std::vector<double> box0;     // R x 3 (box at creation)
    md_options opt{};
    bool skin_set = false;
    std::map<std::pair<int, int>, std::pair<double, double>> pot_lj;  // sigma, eps
    std::map<std::pair<int, int>, std::pair<double, double>> pot_cut; // r_cut, u_cut
    std::vector<PairParam> pairs;
    // layout
    int bs = 256;
    int64_t n_slots = 0, n_atoms = 0, total_cells = 0;
    int n_blocks = 0;
    std::vector<GroupInfo> groups;
    std::vector<RepInfo> reps;
    std::vector<int32_t> blk_group;
    std::vector<int32_t> bld_rep;
    // device
    DevState d{};
    std::vector<void*> base_allocs;
    std::vector<void*> list_allocs;
    void* snap = nullptr;
    size_t snap_bytes = 0;
    RepCtrl* snap_ctrl = nullptr;
    bool fp32 = true, list_mode = true;
    bool sums_valid = false;
    double last_dt = 0.002;
    double rlist = 0.0, rc_max = 0.0;
    int par = 0;       // control buffer the next launch reads
    int snap_par = 0;  // ... at the snapshot
    int n_cu = 256;
    // surrogate
    bool mlp = false, mlp_ready = false;
    std::vector<float> mlp_theta = std::vector<float>(kMlpNP, 0.0f);
    MlpConst mlp_const{};
    int tpb = 1;
    // graphs, by the parity they start at
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int g_len[2] = {0, 0};
    int g_dense[2] = {0, 0};  // the kernel instances (DevState::dense) a graph holds
    // profiling: per-launch device timestamps
    bool prof = false;
    unsigned long long* prof_ring = nullptr;
    unsigned long long* prof_blk = nullptr;
    unsigned long long* prof_host = nullptr;  // pinned staging of the ring
    double wall_hz = 1e8;
    double ktime[P_NUM] = {};      // stream time (previous kernel end to end)
    int64_t kcount[P_NUM] = {};
    double kspan_time[P_NUM] = {};  // first block start to last block end
    int64_t kspan_count[P_NUM] = {};
    int64_t list_visits = 0;
    md_stats stats{};
    int64_t last_rebuilds = 0, last_refresh = 0, last_splits = 0;
    bool ctrl_valid = false;  // ctrl holds epochs of fills into the current grid buffers
    // the step records still hold every replica's state as (x, v + F dt / 2m)
    // for step length crec_dt: the next call needs no pack
    bool crec_valid = false;
    double crec_dt = 0.0;
    // pinned staging of every control block, statistics and parameter copy
    // of the stepping loop (no pageable transfer there)
    RepCtrl* h_ctrl = nullptr;
    RepStats* h_stats = nullptr;
    RepParams* h_params = nullptr;
    RepCtrl* h_ctrl_dev = nullptr;    // device addresses of h_ctrl / h_stats (the epilogue writes there)
    RepStats* h_stats_dev = nullptr;
    hipEvent_t params_ev = nullptr;  // the last parameter upload (h_params is reused after it)
    bool params_pending = false;
    // host copy of the control blocks the next launch reads, valid while no
    // finalize, graph replay, snapshot restore or layout change has run since
    std::vector<RepCtrl> ctrl_cache;
    bool ctrl_cache_valid = false;
    std::string err;
};

// ------------------------------------------------------------------ helpers

static int fail(md_ctx* c, int code, const std::string& msg) {
    if (c)
        c->err = msg;
    else
        g_error = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((c), MD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static double fold_mass(double m, int64_t n) {
    // macro_parameters/mod.rs:17-22: sequential fold of 1.0 * m
    double a = 0.0;
    for (int64_t i = 0; i < n; ++i) a = a + 1.0 * m;
    return a;
}

static void lj_scalar(double sigma, double eps, double r_cut, double u_cut, double r, double* u,
                      double* f) {
    // potential.rs:57-75
    if (r > r_cut) {
        *u = 0.0;
        *f = 0.0;
        return;
    }
    const double sr = sigma / r;
    const double s2 = sr * sr;
    const double s4 = s2 * s2;
    const double s6 = s2 * s4;
    const double s12 = s6 * s6;
    *u = 4.0 * eps * (s12 - s6) - u_cut;
    *f = (24.0 * eps / r) * (s6 - 2.0 * s12);
}

static double r2max_of(double r_cut) {
    // largest double d with sqrt(d) <= r_cut (sqrt is correctly rounded, hence monotone)
    double r2 = r_cut * r_cut;
    while (r2 > 0.0 && sqrt(r2) > r_cut) r2 = nextafter(r2, 0.0);
    while (sqrt(nextafter(r2, INFINITY)) <= r_cut) r2 = nextafter(r2, INFINITY);
    return r2;
}

static void rebuild_pairs(md_ctx* c) {
    const int T = c->T;
    c->pairs.assign((size_t)T * T, PairParam{});
    c->rc_max = 0.0;
    for (int a = 0; a < T; ++a)
        for (int b = 0; b < T; ++b) {
            const std::pair<int, int> key(std::min(a, b), std::max(a, b));
            double sigma = 0.3418, eps = 1.712, rc, uc;  // potential.rs:96
            auto it = c->pot_lj.find(key);
            if (it != c->pot_lj.end()) {
                sigma = it->second.first;
                eps = it->second.second;
            }
            auto ic = c->pot_cut.find(key);
            if (ic != c->pot_cut.end()) {
                rc = ic->second.first;
                uc = ic->second.second;
            } else {
                md_new_lennard_jones(sigma, eps, &rc, &uc);
            }
            PairParam p;
            p.sigma = sigma;
            p.eps = eps;
            p.r_cut = rc;
            p.u_cut = uc;
            p.r2max = r2max_of(rc);
            p.sig2f = (float)(sigma * sigma);
            p.eps24f = (float)(24.0 * eps);
            p.eps4f = (float)(4.0 * eps);
            p.ucutf = (float)uc;
            c->pairs[(size_t)a * T + b] = p;
            c->rc_max = std::max(c->rc_max, rc);
        }
    if (!c->skin_set) c->opt.skin = 0.3 * c->pairs[0].sigma;
    c->rlist = c->rc_max + c->opt.skin;
}

static int dalloc(md_ctx* c, std::vector<void*>& owner, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, MD_ERR_NOMEM, std::string("hipMalloc(") + std::to_string(bytes) +
                                         "): " + hipGetErrorString(e));
    }
    owner.push_back(*p);
    return MD_OK;
}

static void free_list(std::vector<void*>& v) {
    for (void* p : v) (void)hipFree(p);
    v.clear();
}

static void drop_graph(md_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
        c->gexec[k] = nullptr;
        c->g_len[k] = 0;
    }
}

static void free_all(md_ctx* c) {
    drop_graph(c);
    free_list(c->list_allocs);
    free_list(c->base_allocs);
    if (c->snap) (void)hipFree(c->snap);
    c->snap = nullptr;
    c->snap_bytes = 0;
    if (c->snap_ctrl) (void)hipFree(c->snap_ctrl);
    c->snap_ctrl = nullptr;
    if (c->prof_ring) (void)hipFree(c->prof_ring);
    c->prof_ring = nullptr;
    if (c->prof_blk) (void)hipFree(c->prof_blk);
    c->prof_blk = nullptr;
    if (c->prof_host) (void)hipHostFree(c->prof_host);
    c->prof_host = nullptr;
    if (c->h_ctrl) (void)hipHostFree(c->h_ctrl);
    c->h_ctrl = nullptr;
    if (c->h_stats) (void)hipHostFree(c->h_stats);
    c->h_stats = nullptr;
    if (c->params_ev) {
        if (c->params_pending) (void)hipEventSynchronize(c->params_ev);
        (void)hipEventDestroy(c->params_ev);
    }
    c->params_ev = nullptr;
    c->params_pending = false;
    if (c->h_params) (void)hipHostFree(c->h_params);
    c->h_params = nullptr;
    c->h_ctrl_dev = nullptr;
    c->h_stats_dev = nullptr;
    c->d.prof = nullptr;
    c->d.prof_blk = nullptr;
    c->prof = false;  // a new layout starts unprofiled
}

// Tiles per block of the per-slot kernels: the option, else 1 on the
// all-pairs path (compute-bound, wants every block it can get) and on the
// list paths about four blocks per CU, at most 16 tiles each.
static int tiles_per_block(md_ctx* c) {
    if (c->opt.tiles_per_block > 0) return c->opt.tiles_per_block;
    if (c->opt.force_mode == MD_FORCE_FP64_ALLPAIRS) return 1;
    int64_t tiles = 0;
    for (int r = 0; r < c->R; ++r)
        for (int t = 0; t < c->T; ++t) tiles += (c->counts[(size_t)r * c->T + t] + c->bs - 1) / c->bs;
    const int64_t want = 4 * (int64_t)c->n_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(16, (tiles + want - 1) / want));
}

static void make_layout(md_ctx* c) {
    const int bs = c->bs;
    c->tpb = tiles_per_block(c);
    c->groups.clear();
    c->reps.clear();
    c->blk_group.clear();
    int64_t slot = 0, cell_base = 0;
    int block = 0;
    c->n_atoms = 0;
    for (int r = 0; r < c->R; ++r) {
        RepInfo R{};
        R.start = slot;
        R.first_group = (int32_t)c->groups.size();
        R.n_types = c->T;
        R.n_atoms = 0;
        for (int t = 0; t < c->T; ++t) {
            GroupInfo G{};
            G.start = slot;
            G.count = c->counts[(size_t)r * c->T + t];
            const int64_t tiles = (G.count + bs - 1) / bs;
            G.nblocks = (int32_t)((tiles + c->tpb - 1) / c->tpb);
            G.first_block = block;
            G.rep = r;
            G.type = t;
            G.mass = c->masses[t];
            G.sum_m = fold_mass(G.mass, G.count);
            for (int b = 0; b < G.nblocks; ++b) c->blk_group.push_back((int32_t)c->groups.size());
            slot += tiles * bs;
            block += G.nblocks;
            R.n_atoms += G.count;
            c->groups.push_back(G);
        }
        R.end = slot;
        R.first_block = c->groups[R.first_group].first_block;
        R.nblocks = block - R.first_block;
        const double want = ceil(std::max(0.1, c->opt.cells_per_atom) * (double)R.n_atoms);
        R.cell_budget = std::max<int64_t>(27, (int64_t)want);
        R.cell_base = cell_base;
        cell_base += R.cell_budget;
        c->n_atoms += R.n_atoms;
        c->reps.push_back(R);
    }
    c->n_slots = slot;
    c->n_blocks = block;
    c->total_cells = cell_base;
    // list-build blocks: one per CU in total (one 512-thread block fits a CU), shared out by atom count
    int64_t tiles = 0;
    for (const GroupInfo& G : c->groups) tiles += (G.count + bs - 1) / bs;
    const int want = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)c->n_cu));
    c->bld_rep.clear();
    for (int r = 0; r < c->R; ++r) {
        RepInfo& R = c->reps[r];
        const double share = c->n_atoms ? (double)R.n_atoms / (double)c->n_atoms : 1.0 / c->R;
        R.bld_first = (int32_t)c->bld_rep.size();
        R.bld_count = std::max(1, (int)(share * want));
        for (int k = 0; k < R.bld_count; ++k) c->bld_rep.push_back(r);
    }
}

static int auto_capacities(md_ctx* c) {
    // expected neighbours inside r_list and atoms per cell at the current density
    int lc = 8, cc = 8;
    for (int r = 0; r < c->R; ++r) {
        const double* b = &c->box0[(size_t)r * 3];
        const double vol = b[0] * b[1] * b[2];
        const double n = (double)c->reps[r].n_atoms;
        if (n <= 0 || !(vol > 0)) continue;
        const double rho = n / vol;
        const double nb = 4.0 / 3.0 * M_PI * c->rlist * c->rlist * c->rlist * rho;
        lc = std::max(lc, (int)ceil(2.0 * nb + 8.0));
        int gd[3];
        grid_dims(b, c->rlist, c->reps[r].cell_budget, gd);
        const double per = n / ((double)gd[0] * gd[1] * gd[2]);
        cc = std::max(cc, (int)ceil(per + 6.0 * sqrt(per) + 6.0));
    }
    if (c->opt.list_capacity <= 0) c->d.list_cap = lc;
    if (c->opt.cell_capacity <= 0) c->d.cell_cap = cc;
    return MD_OK;
}

static int alloc_lists(md_ctx* c) {
    free_list(c->list_allocs);
    drop_graph(c);
    // the cleared headers match no epoch; the control blocks keep counting
    // theirs, so a later upload still starts past every epoch in use
    void* p;
    int rc;
    c->d.mlp_nbr = nullptr;
    c->d.mlp_ncount = nullptr;
    if (!c->list_mode) {
        c->d.nbr = nullptr;
        c->d.ncount = nullptr;
        c->d.cell_hdr[0] = c->d.cell_hdr[1] = nullptr;
        c->d.cell_slots[0] = c->d.cell_slots[1] = nullptr;
        return MD_OK;
    }
    const int64_t N = c->n_slots;
    if ((rc = dalloc(c, c->list_allocs, &p, sizeof(int32_t) * (size_t)N * c->d.list_cap))) return rc;
    c->d.nbr = (int32_t*)p;
    if ((rc = dalloc(c, c->list_allocs, &p, sizeof(int32_t) * (size_t)N))) return rc;
    c->d.ncount = (int32_t*)p;
    HIPCHK(c, hipMemsetAsync(c->d.ncount, 0, sizeof(int32_t) * (size_t)N, c->st));
    if (c->mlp) {
        // the surrogate's per-step lists of the entries within its cutoff
        if ((rc = dalloc(c, c->list_allocs, &p, sizeof(int32_t) * (size_t)N * c->d.list_cap))) return rc;
        c->d.mlp_nbr = (int32_t*)p;
        if ((rc = dalloc(c, c->list_allocs, &p, sizeof(int32_t) * (size_t)N))) return rc;
        c->d.mlp_ncount = (int32_t*)p;
    }
    c->d.cell_plane = c->total_cells;  // slot k of cell c at k * total_cells + c
    for (int k = 0; k < 2; ++k) {
        if ((rc = dalloc(c, c->list_allocs, &p, sizeof(unsigned long long) * (size_t)c->total_cells)))
            return rc;
        c->d.cell_hdr[k] = (unsigned long long*)p;
        HIPCHK(c, hipMemsetAsync(p, 0, sizeof(unsigned long long) * (size_t)c->total_cells, c->st));
        if ((rc = dalloc(c, c->list_allocs, &p,
                         sizeof(int32_t) * (size_t)c->total_cells * c->d.cell_cap)))
            return rc;
        c->d.cell_slots[k] = (int32_t*)p;
    }
    return MD_OK;
}

static int upload_pairs(md_ctx* c) {
    HIPCHK(c, hipMemcpyAsync((void*)c->d.pairs, c->pairs.data(), sizeof(PairParam) * c->pairs.size(),
                             hipMemcpyHostToDevice, c->st));
    return MD_OK;
}

static bool fp32_records(int mode) { return mode == MD_FORCE_FP32_LIST || mode == MD_FORCE_MLP; }

static size_t rec_size(const md_ctx* c) {
    return fp32_records(c->opt.force_mode) ? sizeof(Rec<float>) : sizeof(Rec<double>);
}

// Surrogate buffers (features, dE/dG, energies, statistics, check sums,
// parameters) and the upload of the host copy of the parameters.
static int mlp_upload(md_ctx* c) {
    if (!c->mlp || !c->mlp_ready) return MD_OK;
    HIPCHK(c, hipMemcpyAsync((void*)c->d.mlp_theta, c->mlp_theta.data(), sizeof(float) * kMlpNP,
                             hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipMemcpyAsync((void*)c->d.mlp_c, &c->mlp_const, sizeof(MlpConst), hipMemcpyHostToDevice, c->st));
    HIPCHK(c, launch_mlp_convert(c->d, c->st));
    return MD_OK;
}

static int alloc_mlp(md_ctx* c) {
    if (!c->mlp) return MD_OK;
    DevState& d = c->d;
    const size_t N = (size_t)std::max<int64_t>(c->n_slots, 1);
    void* p;
    int rc;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(__bf16) * N * kMlpK))) return rc;
    d.mlp_x = (__bf16*)p;
    HIPCHK(c, hipMemset(p, 0, sizeof(__bf16) * N * kMlpK));
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(float) * N * kMlpK))) return rc;
    d.mlp_gam = (float*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(__bf16) * N * kMlpK))) return rc;
    d.mlp_gamb = (__bf16*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(float4) * N))) return rc;
    d.mlp_pos = (float4*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(float) * N))) return rc;
    d.mlp_e = (float*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(MlpStat) * c->R))) return rc;
    d.mlp_stat = (MlpStat*)p;
    HIPCHK(c, hipMemset(p, 0, sizeof(MlpStat) * c->R));
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(double) * 2 * (size_t)std::max(c->n_blocks, 1)))) return rc;
    d.mlp_judge = (double*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(MlpConst)))) return rc;
    d.mlp_c = (const MlpConst*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(float) * kMlpNP))) return rc;
    d.mlp_theta = (const float*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(__bf16) * kMlpBfN))) return rc;
    d.mlp_w = (const __bf16*)p;
    return mlp_upload(c);
}

// Allocate everything for the current layout; state buffers are NaN / zero.
static int alloc_base(md_ctx* c) {
    free_all(c);
    c->crec_valid = false;
    c->ctrl_cache_valid = false;
    make_layout(c);
    c->fp32 = fp32_records(c->opt.force_mode);
    c->list_mode = c->opt.force_mode != MD_FORCE_FP64_ALLPAIRS;
    c->mlp = c->opt.force_mode == MD_FORCE_MLP;
    DevState& d = c->d;
    d = DevState{};
    c->ctrl_valid = false;  // new control blocks: no state, no epochs yet
    const int64_t N = c->n_slots;
    const size_t nd = sizeof(double) * (size_t)std::max<int64_t>(N, 1);
    void* p;
    int rc;
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(c, c->base_allocs, &p, rec_size(c) * (size_t)std::max<int64_t>(N, 1)))) return rc;
        d.rec[b] = p;
    }
    if (c->list_mode && !c->mlp) {
        // the list path steps on position + kicked velocity records
        for (int b = 0; b < 2; ++b) {
            if ((rc = dalloc(c, c->base_allocs, &p, sizeof(CRec) * (size_t)std::max<int64_t>(N, 1)))) return rc;
            d.crec[b] = (CRec*)p;
        }
        if ((rc = dalloc(c, c->base_allocs, &p, sizeof(double4) * (size_t)std::max<int64_t>(N, 1)))) return rc;
        d.dpos = (double4*)p;
        std::vector<uint8_t> ty((size_t)std::max<int64_t>(N, 1), 0);
        for (const GroupInfo& G : c->groups) {
            const int64_t n = (G.count + c->bs - 1) / c->bs * c->bs;
            for (int64_t i = 0; i < n; ++i) ty[(size_t)(G.start + i)] = (uint8_t)G.type;
        }
        if ((rc = dalloc(c, c->base_allocs, &p, ty.size()))) return rc;
        HIPCHK(c, hipMemcpy(p, ty.data(), ty.size(), hipMemcpyHostToDevice));
        d.slot_type = (const uint8_t*)p;
        d.compact = 1;
    }
    if ((rc = dalloc(c, c->base_allocs, &p, nd))) return rc;
    d.pot = (double*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, nd))) return rc;
    d.vir = (double*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(int32_t) * std::max<size_t>(c->blk_group.size(), 1))))
        return rc;
    d.blk_group = (const int32_t*)p;
    if (!c->blk_group.empty())
        HIPCHK(c, hipMemcpy(p, c->blk_group.data(), sizeof(int32_t) * c->blk_group.size(),
                            hipMemcpyHostToDevice));
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(int32_t) * c->bld_rep.size()))) return rc;
    d.bld_rep = (const int32_t*)p;
    HIPCHK(c, hipMemcpy(p, c->bld_rep.data(), sizeof(int32_t) * c->bld_rep.size(), hipMemcpyHostToDevice));
    d.n_bld = (int32_t)c->bld_rep.size();
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(GroupInfo) * c->groups.size()))) return rc;
    d.groups = (const GroupInfo*)p;
    HIPCHK(c, hipMemcpy(p, c->groups.data(), sizeof(GroupInfo) * c->groups.size(),
                        hipMemcpyHostToDevice));
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(RepInfo) * c->reps.size()))) return rc;
    d.reps = (const RepInfo*)p;
    HIPCHK(c, hipMemcpy(p, c->reps.data(), sizeof(RepInfo) * c->reps.size(), hipMemcpyHostToDevice));
    if ((rc = dalloc(c, c->base_allocs, &p, 2 * sizeof(RepCtrl) * c->R))) return rc;
    d.ctrl = (RepCtrl*)p;
    HIPCHK(c, hipMemset(p, 0, 2 * sizeof(RepCtrl) * c->R));
    c->par = 0;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(RepParams) * c->R))) return rc;
    d.params = (const RepParams*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(RepStats)))) return rc;
    d.stats = (RepStats*)p;
    HIPCHK(c, hipMemset(p, 0, sizeof(RepStats)));
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(PairParam) * (size_t)c->T * c->T))) return rc;
    d.pairs = (const PairParam*)p;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(double) * NS * (size_t)std::max(c->n_blocks, 1))))
        return rc;
    d.partial = (double*)p;
    d.acc_mode = (c->R == 1 && c->T == 1) ? 1 : 0;
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(unsigned long long) * kAccWords))) return rc;
    d.acc = (unsigned long long*)p;
    HIPCHK(c, hipMemset(p, 0, sizeof(unsigned long long) * kAccWords));
    if ((rc = dalloc(c, c->base_allocs, &p, sizeof(double) * NS * c->groups.size()))) return rc;
    d.sums = (double*)p;
    HIPCHK(c, hipMemset(p, 0, sizeof(double) * NS * c->groups.size()));
    d.n_types = c->T;
    d.n_blocks = c->n_blocks;
    d.n_reps = c->R;
    d.tpb = c->tpb;
    d.par = 0;
    d.n_slots = N;
    d.list_cap = c->opt.list_capacity > 0 ? c->opt.list_capacity : 8;
    d.cell_cap = c->opt.cell_capacity > 0 ? c->opt.cell_capacity : 8;
    rebuild_pairs(c);
    if ((rc = upload_pairs(c))) return rc;
    auto_capacities(c);
    if ((rc = alloc_lists(c))) return rc;
    if ((rc = alloc_mlp(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return MD_OK;
}

// Control blocks of the buffer the next launch reads.
static RepCtrl* ctrl_cur(md_ctx* c) { return c->d.ctrl + (size_t)c->par * c->R; }

// Pinned host staging of the stepping loop's transfers (allocated once per
// layout).
static int pinned_staging(md_ctx* c) {
    if (c->h_ctrl) return MD_OK;
    void* p = nullptr;
    HIPCHK(c, hipHostMalloc(&p, sizeof(RepCtrl) * c->R, hipHostMallocDefault));
    c->h_ctrl = (RepCtrl*)p;
    HIPCHK(c, hipHostMalloc(&p, sizeof(RepStats), hipHostMallocDefault));
    c->h_stats = (RepStats*)p;
    HIPCHK(c, hipHostMalloc(&p, sizeof(RepParams) * c->R, hipHostMallocDefault));
    c->h_params = (RepParams*)p;
    HIPCHK(c, hipHostGetDevicePointer(&p, c->h_ctrl, 0));
    c->h_ctrl_dev = (RepCtrl*)p;
    HIPCHK(c, hipHostGetDevicePointer(&p, c->h_stats, 0));
    c->h_stats_dev = (RepStats*)p;
    HIPCHK(c, hipEventCreateWithFlags(&c->params_ev, hipEventDisableTiming));
    return MD_OK;
}

static int read_ctrl(md_ctx* c, std::vector<RepCtrl>& out) {
    if (c->ctrl_cache_valid && (int)c->ctrl_cache.size() == c->R) {
        out = c->ctrl_cache;
        return MD_OK;
    }
    int rc;
    if ((rc = pinned_staging(c))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->h_ctrl, ctrl_cur(c), sizeof(RepCtrl) * c->R, hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    out.assign(c->h_ctrl, c->h_ctrl + c->R);
    c->ctrl_cache = out;
    c->ctrl_cache_valid = true;
    return MD_OK;
}

static int write_ctrl(md_ctx* c, const std::vector<RepCtrl>& in) {
    int rc;
    if ((rc = pinned_staging(c))) return rc;
    // every use of h_ctrl ends in a stream synchronisation, so it is free here
    memcpy(c->h_ctrl, in.data(), sizeof(RepCtrl) * c->R);
    HIPCHK(c, hipMemcpyAsync(ctrl_cur(c), c->h_ctrl, sizeof(RepCtrl) * c->R, hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->ctrl_cache = in;
    c->ctrl_cache_valid = true;
    return MD_OK;
}

// Finalize launch; M_STEP and M_PRIME write the other control buffer.
static hipError_t finalize(md_ctx* c, int mode) {
    if (mode != M_REDUCE) c->ctrl_cache_valid = false;
    c->d.par = c->par;
    hipError_t e = launch_finalize(c->d, mode, c->list_mode, c->fp32, c->st);
    if (e == hipSuccess && mode != M_REDUCE) c->par ^= 1;
    return e;
}

static void set_kick(md_ctx* c, RepParams& P, double dt) {
    P.dt = dt;
    for (int t = 0; t < c->T; ++t) {
        P.cm[t] = dt / (2.0 * c->masses[t]);  // integrator.rs:29 temp = delta_time / (2.0 * mass)
        P.inv_m2[t] = 1.0 / (c->masses[t] * c->masses[t]);
    }
}

static RepParams base_params(md_ctx* c) {
    RepParams P{};
    set_kick(c, P, c->last_dt);
    P.list_mode = c->list_mode ? 1 : 0;
    P.in_place = c->mlp ? 1 : 0;
    P.rlist = c->rlist;
    P.rc_max = c->rc_max;
    P.margin = 1e-9 * c->rlist;
    P.insert_growth = 1.0;
    P.target_step = 0;
    P.full_dst = -1;
    return P;
}

static int write_params(md_ctx* c, const std::vector<RepParams>& P) {
    int rc;
    if ((rc = pinned_staging(c))) return rc;
    // the previous upload may still read the staging buffer
    if (c->params_pending) HIPCHK(c, hipEventSynchronize(c->params_ev));
    memcpy(c->h_params, P.data(), sizeof(RepParams) * c->R);
    HIPCHK(c, hipMemcpyAsync((void*)c->d.params, c->h_params, sizeof(RepParams) * c->R,
                             hipMemcpyHostToDevice, c->st));
    HIPCHK(c, hipEventRecord(c->params_ev, c->st));
    c->params_pending = true;
    return MD_OK;
}

// Group sums of the current state.  The thermal partials are taken about
// the centre-of-mass velocity of the sums stored before; after a new state
// those belong to another state, so a second pass takes them about the new
// centre of mass (no cancellation whatever the centre of mass does).
static int reduce_state_sums(md_ctx* c, int passes) {
    c->d.par = c->par;
    for (int k = 0; k < passes; ++k) {
        if (c->n_blocks > 0) HIPCHK(c, launch_reduce_state(c->d, c->fp32, c->bs, c->st));
        HIPCHK(c, finalize(c, M_REDUCE));
    }
    return MD_OK;
}

static int ensure_sums(md_ctx* c) {
    if (c->sums_valid) return MD_OK;
    std::vector<RepParams> P((size_t)c->R, base_params(c));
    int rc;
    if ((rc = write_params(c, P))) return rc;
    if ((rc = reduce_state_sums(c, 2))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->sums_valid = true;
    return MD_OK;
}

static int read_sums(md_ctx* c, std::vector<double>& S) {
    S.resize((size_t)NS * c->groups.size());
    HIPCHK(c, hipMemcpyAsync(S.data(), c->d.sums, sizeof(double) * S.size(), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return MD_OK;
}

// ------------------------------------------------------------------ profiling

// Fold the device timestamp ring into the per-kernel totals and reset it.
static int prof_collect(md_ctx* c) {
    if (!c->prof || !c->prof_ring) return MD_OK;
    const size_t n = (size_t)kProfCount + 1;
    if (!c->prof_host) {
        void* p = nullptr;
        HIPCHK(c, hipHostMalloc(&p, sizeof(unsigned long long) * n, hipHostMallocDefault));
        c->prof_host = (unsigned long long*)p;
    }
    const unsigned long long* h = c->prof_host;  // pinned: the copy is a DMA transfer
    HIPCHK(c, hipMemcpyAsync(c->prof_host, c->prof_ring, sizeof(unsigned long long) * n,
                             hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    // span: first block start to last block end of each kernel launch
    for (size_t e = 0; e < (size_t)kProfRing * P_NUM; ++e) {
        const unsigned long long a = h[2 * e], b = h[2 * e + 1];
        if (a == ~0ull || b < a) continue;
        const int k = (int)(e % P_NUM);
        c->kspan_time[k] += (double)(b - a) / c->wall_hz * 1e3;
        c->kspan_count[k] += 1;
    }
    // stream time: from the end of the previous recorded kernel on the stream
    // (in launch order drift, features, net, step, finalize) to this kernel's end, counted
    // where the two ran back to back (the interval a dispatch trace reports)
    const unsigned long long gap = (unsigned long long)(20e-6 * c->wall_hz);
    auto rec = [&](int l, int k, unsigned long long* a, unsigned long long* b) {
        const size_t e = ((size_t)l * P_NUM + k) * 2;
        *a = h[e];
        *b = h[e + 1];
        return *a != ~0ull && *b >= *a;
    };
    for (int l = 0; l < kProfRing; ++l) {
        unsigned long long pa, pb, a, b;
        const int lp = (l + kProfRing - 1) % kProfRing;
        bool have_prev = rec(lp, P_FIN, &pa, &pb);
        unsigned long long prev_end = pb;
        for (int k : {(int)P_DRIFT, (int)P_FEAT, (int)P_NET, (int)P_STEP, (int)P_FIN}) {
            if (!rec(l, k, &a, &b)) continue;
            if (have_prev && a >= prev_end && a - prev_end < gap) {
                c->ktime[k] += (double)(b - prev_end) / c->wall_hz * 1e3;
                c->kcount[k] += 1;
            }
            have_prev = true;
            prev_end = b;
        }
    }
    c->list_visits += (int64_t)h[(size_t)kProfCount];
    HIPCHK(c, launch_prof_init(c->d, c->st));
    return MD_OK;
}

// One step launch: list paths k_step + k_fin (which also builds lists),
// all-pairs path k_drift + k_allpairs + k_fin.
static hipError_t launch_once(md_ctx* c) {
    hipError_t e;
    c->d.par = c->par;
    if (c->mlp) {
        if ((e = launch_mlp_desc(c->d, c->bs, false, 1, c->st))) return e;
        if ((e = launch_mlp_desc(c->d, c->bs, false, 2, c->st))) return e;
        if ((e = launch_mlp_net(c->d, c->n_cu, c->st))) return e;
        if ((e = launch_mlp_force(c->d, c->bs, false, c->st))) return e;
    } else if (c->list_mode) {
        if (c->d.dense && (e = launch_dpos(c->d, c->bs, c->st))) return e;
        if ((e = launch_step(c->d, c->fp32, c->bs, c->st))) return e;
    } else {
        if ((e = launch_drift(c->d, c->bs, c->st))) return e;
        if ((e = launch_allpairs(c->d, c->bs, false, c->st))) return e;
    }
    return finalize(c, M_STEP);
}

// Graph of `len` launches starting at the current parity (len is even, so a
// replay ends at the parity it started at).
static int ensure_graph(md_ctx* c, int len) {
    const int p0 = c->par;
    if (c->gexec[p0] && c->g_len[p0] == len && c->g_dense[p0] == c->d.dense) return MD_OK;
    if (c->gexec[p0]) (void)hipGraphExecDestroy(c->gexec[p0]);
    c->gexec[p0] = nullptr;
    c->g_len[p0] = 0;
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < len && e == hipSuccess; ++k) e = launch_once(c);
    hipError_t e2 = hipStreamEndCapture(c->st, &g);
    c->par = p0;
    if (e != hipSuccess) return fail(c, MD_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
    if (e2 != hipSuccess) return fail(c, MD_ERR_HIP, std::string("end capture: ") + hipGetErrorString(e2));
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, MD_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    c->gexec[p0] = x;
    c->g_len[p0] = len;
    c->g_dense[p0] = c->d.dense;
    return MD_OK;
}

// ------------------------------------------------------------------ snapshot

// The step-record path leaves each replica's starting buffer untouched
// during a chunk (its last step writes the other one), so its snapshot
// needs no records.
static bool records_in_snapshot(const md_ctx* c) { return !c->d.compact; }

static size_t state_bytes(md_ctx* c) {
    // both record buffers (replicas after a split step sit at different
    // parities), pot, vir, the group sums the first step's controls derive
    // from, and the surrogate's counters
    const size_t recs = records_in_snapshot(c) ? (2 * rec_size(c) + 2 * sizeof(double)) * (size_t)c->n_slots : 0;
    return recs + sizeof(double) * NS * c->groups.size() + (c->d.mlp_stat ? sizeof(MlpStat) * c->R : 0);
}

static int snapshot_alloc(md_ctx* c) {
    if (!c->snap) {
        c->snap_bytes = std::max<size_t>(state_bytes(c), 256);
        HIPCHK(c, hipMalloc(&c->snap, c->snap_bytes));
        HIPCHK(c, hipMalloc((void**)&c->snap_ctrl, sizeof(RepCtrl) * c->R));
    }
    return MD_OK;
}

static int snapshot(md_ctx* c, bool save) {
    const size_t n = sizeof(double) * (size_t)c->n_slots;
    int rc;
    if ((rc = snapshot_alloc(c))) return rc;
    char* s = (char*)c->snap;
    const bool recs = records_in_snapshot(c);
    void* arrs[6] = {c->d.rec[0], c->d.rec[1], c->d.pot, c->d.vir, c->d.sums, c->d.mlp_stat};
    const size_t rs = recs ? rec_size(c) * (size_t)c->n_slots : 0, ps = recs ? n : 0;
    const size_t sz[6] = {rs, rs, ps, ps, sizeof(double) * NS * c->groups.size(),
                          c->d.mlp_stat ? sizeof(MlpStat) * c->R : 0};
    size_t off = 0;
    for (int k = 0; k < 6; ++k) {
        if (sz[k] == 0) continue;
        if (save)
            HIPCHK(c, hipMemcpyAsync(s + off, arrs[k], sz[k], hipMemcpyDeviceToDevice, c->st));
        else
            HIPCHK(c, hipMemcpyAsync(arrs[k], s + off, sz[k], hipMemcpyDeviceToDevice, c->st));
        off += sz[k];
    }
    if (save) {
        c->snap_par = c->par;
        HIPCHK(c, hipMemcpyAsync(c->snap_ctrl, ctrl_cur(c), sizeof(RepCtrl) * c->R, hipMemcpyDeviceToDevice, c->st));
    } else {
        c->par = c->snap_par;
        c->ctrl_cache_valid = false;
        HIPCHK(c, hipMemcpyAsync(ctrl_cur(c), c->snap_ctrl, sizeof(RepCtrl) * c->R, hipMemcpyDeviceToDevice, c->st));
    }
    return MD_OK;
}

static int reset_stats(md_ctx* c) {
    HIPCHK(c, hipMemsetAsync(c->d.stats, 0, sizeof(RepStats), c->st));
    return MD_OK;
}

static int read_stats(md_ctx* c, RepStats& s) {
    HIPCHK(c, hipMemcpyAsync(&s, c->d.stats, sizeof(RepStats), hipMemcpyDeviceToHost, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return MD_OK;
}

// Control blocks and statistics after a chunk, with one wait.
static int read_ctrl_stats(md_ctx* c, std::vector<RepCtrl>& C, RepStats& s) {
    int rc;
    if ((rc = pinned_staging(c))) return rc;
    // one kernel writes both into the pinned buffers (two copies would be
    // two device operations)
    HIPCHK(c, launch_epilogue(c->d, c->par, c->h_ctrl_dev, c->h_stats_dev, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    C.assign(c->h_ctrl, c->h_ctrl + c->R);
    s = *c->h_stats;
    c->ctrl_cache = C;
    c->ctrl_cache_valid = true;
    return MD_OK;
}

// Grow list / cell capacity; returns 1 when grown.  After an overflow (the
// chunk is replayed) the capacity at least doubles; ahead of one (a list or
// cell within a quarter of its capacity) it grows by a quarter: dense
// builds keep each centre particle's list column in a fixed LDS pool, so a
// capacity far above the longest list costs build passes.
static int grow_capacity(md_ctx* c, const RepStats& s, int* grown, bool ahead = false) {
    *grown = 0;
    if (s.list_overflow) {
        const int cap = c->d.list_cap;
        c->d.list_cap = std::max(ahead ? cap + (cap + 3) / 4 : cap * 2, (int)(s.max_list * 1.25) + 4);
        *grown = 1;
    }
    if (s.cell_overflow) {
        const int cap = c->d.cell_cap;
        c->d.cell_cap = std::max(ahead ? cap + (cap + 3) / 4 : cap * 2, (int)(s.max_cell * 1.25) + 4);
        *grown = 1;
    }
    if (*grown) return alloc_lists(c);
    return MD_OK;
}

// Build lists at the current positions and evaluate forces (list paths).
static int list_force_eval(md_ctx* c) {
    int rc;
    for (int attempt = 0; attempt < 8; ++attempt) {
        std::vector<RepCtrl> C;
        if ((rc = read_ctrl(c, C))) return rc;
        for (int r = 0; r < c->R; ++r) {
            RepCtrl& X = C[r];
            X.ins_epoch = X.fill_epoch + 1u;
            const int b = X.ins_epoch & 1;
            for (int k = 0; k < 3; ++k) X.gbox[b][k] = X.box[k];
            grid_dims(X.gbox[b], c->rlist, c->reps[r].cell_budget, X.gdim[b]);
            brick_dims(X.gdim[b], c->reps[r].n_atoms, X.bedge[b], &X.bwhole[b]);
            X.fill_epoch = X.ins_epoch;
            X.build_src = X.cur;
            X.flags = F_BUILD;
        }
        if ((rc = write_ctrl(c, C))) return rc;
        std::vector<RepParams> P((size_t)c->R, base_params(c));
        if ((rc = write_params(c, P))) return rc;
        if ((rc = reset_stats(c))) return rc;
        c->d.par = c->par;
        if (c->n_blocks > 0) {
            HIPCHK(c, launch_insert_all(c->d, c->fp32, c->bs, c->st));
            HIPCHK(c, launch_build(c->d, c->fp32, c->st));
            if (c->mlp) {
                HIPCHK(c, launch_mlp_desc(c->d, c->bs, true, 0, c->st));
                HIPCHK(c, launch_mlp_net(c->d, c->n_cu, c->st));
                HIPCHK(c, launch_mlp_force(c->d, c->bs, true, c->st));
            } else {
                HIPCHK(c, launch_force_eval(c->d, c->fp32, c->bs, c->st));
            }
        }
        HIPCHK(c, finalize(c, M_REDUCE));
        if ((rc = reduce_state_sums(c, 1))) return rc;
        RepStats s;
        if ((rc = read_stats(c, s))) return rc;
        int grown = 0;
        if ((rc = grow_capacity(c, s, &grown))) return rc;
        if (grown) continue;
        if ((rc = read_ctrl(c, C))) return rc;
        for (int r = 0; r < c->R; ++r) {
            C[r].flags = F_HAVE_LIST | F_GRID_FRESH;
            C[r].D = 0.0;
            C[r].s = 1.0;
        }
        if ((rc = write_ctrl(c, C))) return rc;
        c->sums_valid = true;
        c->stats.list_capacity = c->d.list_cap;
        c->stats.cell_capacity = c->d.cell_cap;
        return MD_OK;
    }
    return fail(c, MD_ERR_OVERFLOW, "neighbour list / cell capacity kept overflowing");
}

// ------------------------------------------------------------------ step

static int run_steps(md_ctx* c, double dt, md_thermostat* th, md_barostat* ba, int64_t n,
                     bool per_rep) {
    if (n < 0) return fail(c, MD_ERR_ARG, "n_steps < 0");
    if (n == 0) return MD_OK;
    if (c->mlp && !c->mlp_ready) return fail(c, MD_ERR_ARG, "surrogate parameters not set (md_mlp_set_params)");
    for (int r = 0; r < c->R; ++r) {
        const md_thermostat* t = th ? &th[per_rep ? r : 0] : nullptr;
        const md_barostat* b = ba ? &ba[per_rep ? r : 0] : nullptr;
        if (t && (t->kind == 3)) return fail(c, MD_ERR_UNIMPLEMENTED, "Thermostat::Custom is todo!() in the reference");
        if (t && (t->kind < 0 || t->kind > 3)) return fail(c, MD_ERR_ARG, "bad thermostat kind");
        if (b && (b->kind == 2)) return fail(c, MD_ERR_UNIMPLEMENTED, "Barostat::Custom is todo!() in the reference");
        if (b && (b->kind < 0 || b->kind > 2)) return fail(c, MD_ERR_ARG, "bad barostat kind");
    }
    for (const GroupInfo& G : c->groups)
        if (G.count == 0)
            return fail(c, MD_ERR_ARG, "a particle type has no particles (the reference indexes particle_type[0])");
    c->last_dt = dt;
    int rc;
    if ((rc = ensure_sums(c))) return rc;
    std::vector<RepCtrl> C;
    if ((rc = read_ctrl(c, C))) return rc;
    const int64_t step0 = C[0].step;
    std::vector<RepParams> P((size_t)c->R, base_params(c));
    for (int r = 0; r < c->R; ++r) {
        RepParams& p = P[r];
        set_kick(c, p, dt);
        const md_thermostat* t = th ? &th[per_rep ? r : 0] : nullptr;
        const md_barostat* b = ba ? &ba[per_rep ? r : 0] : nullptr;
        if (t && t->kind > 0) {
            p.th_kind = t->kind;
            p.th_tau = t->tau;
            p.th_target = t->target_t;
            p.psi_in = t->kind == 2 ? t->psi : 0.0;
        }
        if (b && b->kind > 0) {
            p.ba_kind = b->kind;
            p.ba_beta = b->beta;
            p.ba_tau = b->tau;
            p.ba_target = b->target_p;
        }
        // the box and positions scale inconsistently with several types under a
        // barostat (barostat.rs:45 runs once per type): no skin can be trusted
        p.rebuild_every_step = (c->list_mode && c->T > 1 && p.ba_kind == 1) ? 1 : 0;
        if (C[r].step != step0) return fail(c, MD_ERR_ARG, "replicas out of step");
    }
    // graphs hold an even number of launches so that a replay ends at the
    // control buffer it started from
    int K = c->opt.graph_chunk > 0 ? c->opt.graph_chunk : 32;
    if (K > 1) K += K & 1;
    const int64_t chunk = c->list_mode ? std::max<int64_t>(512, 16 * (int64_t)K) : 64;
    int64_t done = 0;
    bool first = true, bad = false;
    int retries = 0;
    md_stats st{};
    while (done < n) {
        const int64_t cn = std::min(n - done, chunk);
        // one replica: parameters, cleared statistics and (step-record path)
        // the snapshot in one prologue kernel instead of several copies
        const bool one = c->R == 1;
        bool snap_in_prologue = false;
        if (c->list_mode) {
            if (one && !records_in_snapshot(c)) {
                if ((rc = snapshot_alloc(c))) return rc;
                c->snap_par = c->par;
                snap_in_prologue = true;
            } else if ((rc = snapshot(c, true))) {
                return rc;
            }
            // dense systems (long lists expected at the current densities)
            // step and build with the long-list kernel instances
            double ne = 0.0;
            for (int r = 0; r < c->R; ++r) {
                const double v = C[r].box[0] * C[r].box[1] * C[r].box[2];
                if (v > 0.0)
                    ne = std::max(ne, (double)c->reps[r].n_atoms / v * (4.0 / 3.0 * M_PI) * c->rlist * c->rlist *
                                          c->rlist);
            }
            c->d.dense = ne >= (double)kLaneWalk ? 1 : 0;
        }
        for (int r = 0; r < c->R; ++r) {
            P[r].target_step = step0 + done + cn;
            P[r].use_psi_in = first ? 1 : 0;
            P[r].full_dst = c->d.compact ? ((C[r].cur & 1) ^ 1) : -1;
        }
        if (one) {
            // the snapshot's layout without records: the group sums, then the
            // surrogate's counters (snapshot())
            char* sb = (char*)c->snap;
            const size_t sums_bytes = sizeof(double) * NS * c->groups.size();
            HIPCHK(c, launch_prologue(c->d, P[0], snap_in_prologue ? (double*)sb : nullptr, (int)c->groups.size(),
                                      snap_in_prologue && c->d.mlp_stat ? (MlpStat*)(sb + sums_bytes) : nullptr,
                                      snap_in_prologue ? c->snap_ctrl : nullptr, c->par, c->st));
        } else {
            if ((rc = write_params(c, P))) return rc;
            if ((rc = reset_stats(c))) return rc;
        }
        if (c->d.compact && !(c->crec_valid && c->crec_dt == dt)) {
            // the chunk steps on step records; its last step writes the full
            // ones back (and leaves the step records valid for the next call)
            c->d.par = c->par;
            HIPCHK(c, launch_pack(c->d, c->fp32, c->bs, c->st));
        }
        HIPCHK(c, finalize(c, M_PRIME));
        int64_t k = 0;
        if (K > 1 && cn >= K) {
            if ((rc = ensure_graph(c, K))) return rc;
            c->ctrl_cache_valid = false;
            for (; k + K <= cn; k += K) HIPCHK(c, hipGraphLaunch(c->gexec[c->par], c->st));
        }
        for (; k < cn; ++k) HIPCHK(c, launch_once(c));
        int64_t launched = cn;
        RepStats s;
        for (int iter = 0;; ++iter) {
            if ((rc = read_ctrl_stats(c, C, s))) return rc;
            int64_t behind = 0;
            for (int r = 0; r < c->R; ++r) behind = std::max(behind, P[r].target_step - C[r].step);
            if (behind <= 0) break;
            if (iter > 64) return fail(c, MD_ERR_HIP, "replicas did not reach the target step");
            for (int64_t q = 0; q < behind; ++q) HIPCHK(c, launch_once(c));
            launched += behind;
        }
        c->crec_valid = false;  // set again once the chunk stands
        if (s.list_overflow || s.cell_overflow) {
            int grown = 0;
            if ((rc = grow_capacity(c, s, &grown))) return rc;
            if (++retries > 8) return fail(c, MD_ERR_OVERFLOW, "capacity kept overflowing");
            if ((rc = snapshot(c, false))) return rc;
            if ((rc = read_ctrl(c, C))) return rc;
            for (int r = 0; r < c->R; ++r) C[r].flags &= ~(F_HAVE_LIST | F_GRID_FRESH);
            if ((rc = write_ctrl(c, C))) return rc;
            if (c->prof) HIPCHK(c, launch_prof_init(c->d, c->st));  // the replay reuses launch indices
            st.retries++;
            continue;
        }
        if ((rc = prof_collect(c))) return rc;
        // a list or cell that came within a quarter of its capacity: grow now,
        // lists are rebuilt at the next launch
        if (s.max_list * 4 > c->d.list_cap * 3 || s.max_cell * 4 > c->d.cell_cap * 3) {
            RepStats g{};
            g.list_overflow = s.max_list * 4 > c->d.list_cap * 3;
            g.cell_overflow = s.max_cell * 4 > c->d.cell_cap * 3;
            g.max_list = s.max_list;
            g.max_cell = s.max_cell;
            int grown = 0;
            if ((rc = grow_capacity(c, g, &grown, true))) return rc;
            for (int r = 0; r < c->R; ++r) C[r].flags &= ~(F_HAVE_LIST | F_GRID_FRESH);
            if ((rc = write_ctrl(c, C))) return rc;
            st.grows++;
        }
        for (int r = 0; r < c->R; ++r)
            if (C[r].flags & F_BAD) bad = true;
        c->crec_valid = c->d.compact != 0;
        c->crec_dt = dt;
        st.launches += launched;
        done += cn;
        first = false;
    }
    for (int r = 0; r < c->R; ++r) {
        md_thermostat* t = th ? &th[per_rep ? r : 0] : nullptr;
        md_barostat* b = ba ? &ba[per_rep ? r : 0] : nullptr;
        if (t && t->kind > 0 && (per_rep || r == 0)) {
            t->lambda = C[r].lambda_used;
            if (t->kind == 2) t->psi = C[r].psi_used;
        }
        if (b && b->kind > 0 && (per_rep || r == 0)) b->myu = C[r].myu_used;
    }
    st.steps = n;
    int64_t rb = 0, nl = 0, ns = 0;
    for (int r = 0; r < c->R; ++r) {
        rb += C[r].rebuilds;
        nl += C[r].nulls;
        ns += C[r].splits;
    }
    st.rebuilds = rb - c->stats.rebuilds;  // counters in ctrl are cumulative
    st.refresh_steps = nl - c->stats.refresh_steps;
    st.split_steps = ns - c->stats.split_steps;
    c->stats.rebuilds = rb;
    c->stats.refresh_steps = nl;
    c->stats.split_steps = ns;
    c->stats.steps = st.steps;
    c->stats.launches = st.launches;
    c->stats.retries = st.retries;
    c->stats.grows = st.grows;
    c->stats.list_capacity = c->d.list_cap;
    c->stats.cell_capacity = c->d.cell_cap;
    c->stats.n_slots = (int32_t)c->n_slots;
    c->last_rebuilds = st.rebuilds;
    c->last_refresh = st.refresh_steps;
    c->last_splits = st.split_steps;
    c->sums_valid = true;
    if (bad) return fail(c, MD_ERR_BAROSTAT, "non-finite temperature/pressure or mu^3 <= 0 during the run");
    return MD_OK;
}

// ------------------------------------------------------------------ C ABI

// ------------------------------------------------------------------ frame source

void md_set_global_error(const std::string& msg) { g_error = msg; }

int md_frame_source(md_ctx* c, int replica, md::FrameSource* out) {
    if (!c || !out) return MD_ERR_ARG;
    if (replica < 0 || replica >= c->R) return fail(c, MD_ERR_ARG, "bad replica");
    md::FrameSource& s = *out;
    s = md::FrameSource{};
    s.st = c->st;
    s.device = c->device;
    s.rec[0] = (const unsigned char*)c->d.rec[0];
    s.rec[1] = (const unsigned char*)c->d.rec[1];
    s.stride = c->fp32 ? (int32_t)sizeof(Rec<float>) : (int32_t)sizeof(Rec<double>);
    s.ctrl = ctrl_cur(c) + replica;
    int64_t row = 0;
    for (const GroupInfo& G : c->groups) {
        if (G.rep != replica) continue;
        s.row0[s.n_groups] = row;
        s.slot0[s.n_groups] = G.start;
        s.type[s.n_groups] = G.type;
        row += G.count;
        ++s.n_groups;
    }
    s.row0[s.n_groups] = row;
    return MD_OK;
}

extern "C" {

void md_lj_potential_and_force(double sigma, double eps, double r_cut, double u_cut, double r,
                               double* u, double* dudr) {
    double a, b;
    lj_scalar(sigma, eps, r_cut, u_cut, r, &a, &b);
    if (u) *u = a;
    if (dudr) *dudr = b;
}

void md_new_lennard_jones(double sigma, double eps, double* r_cut, double* u_cut) {
    // potential.rs:27-55: r_cut = sigma * 2.5, u_cut = U(r_cut) evaluated with u_cut = 0
    const double rc = sigma * 2.5;
    double u, f;
    lj_scalar(sigma, eps, rc, 0.0, rc, &u, &f);
    if (r_cut) *r_cut = rc;
    if (u_cut) *u_cut = u;
}

const char* md_global_error(void) { return g_error.c_str(); }

const char* md_last_error(const md_ctx* c) { return c ? c->err.c_str() : g_error.c_str(); }

static int check_device(int* dev) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(nullptr, MD_ERR_NO_DEVICE, "no HIP device visible");
    if (hipGetDevice(dev) != hipSuccess) return fail(nullptr, MD_ERR_NO_DEVICE, "hipGetDevice failed");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, *dev) != hipSuccess)
        return fail(nullptr, MD_ERR_NO_DEVICE, "hipGetDeviceProperties failed");
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, MD_ERR_NO_DEVICE,
                    std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
    return MD_OK;
}

int md_create_replicas(md_ctx** out, int n_replicas, int n_types, const int64_t* counts,
                       const double* masses, const double* boxes) {
    if (!out) return fail(nullptr, MD_ERR_ARG, "out is NULL");
    *out = nullptr;
    if (n_replicas < 1 || n_types < 1 || n_types > kMaxTypes || !counts || !masses || !boxes)
        return fail(nullptr, MD_ERR_ARG, "bad replica/type counts or NULL arrays (1..8 types)");
    int64_t total = 0;
    for (int64_t i = 0; i < (int64_t)n_replicas * n_types; ++i) {
        if (counts[i] < 0) return fail(nullptr, MD_ERR_ARG, "negative particle count");
        // each term is checked before the sum so that it cannot overflow
        if (counts[i] >= ((int64_t)1 << 30) - total)
            return fail(nullptr, MD_ERR_ARG, "too many particles for 32-bit slot ids");
        total += counts[i];
    }
    int dev = 0;
    int rc = check_device(&dev);
    if (rc) return rc;
    md_ctx* c = new md_ctx();
    c->device = dev;
    c->R = n_replicas;
    c->T = n_types;
    c->counts.assign(counts, counts + (size_t)n_replicas * n_types);
    c->masses.assign(masses, masses + n_types);
    c->box0.assign(boxes, boxes + (size_t)n_replicas * 3);
    c->opt.force_mode = MD_FORCE_FP32_LIST;
    c->opt.skin = 0.0;
    c->opt.block_size = 256;
    c->opt.list_capacity = 0;
    c->opt.cell_capacity = 0;
    c->opt.cells_per_atom = 2.0;
    c->opt.graph_chunk = 32;
    c->opt.tiles_per_block = 0;
    c->bs = 256;
    {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
            c->n_cu = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeWallClockRate, dev) == hipSuccess && v > 0)
            c->wall_hz = 1e3 * (double)v;  // attribute is in kHz
    }
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, MD_ERR_HIP, "hipStreamCreate failed");
    }
    rc = alloc_base(c);
    if (rc) {
        g_error = c->err;
        md_destroy(c);
        return rc;
    }
    std::vector<double> zeros((size_t)total * 3, 0.0);
    rc = md_upload_state(c, zeros.data(), zeros.data(), c->box0.data());
    if (rc) {
        g_error = c->err;
        md_destroy(c);
        return rc;
    }
    *out = c;
    return MD_OK;
}

int md_create(md_ctx** out, int n_types, const int64_t* counts, const double* masses,
              const double box[3]) {
    return md_create_replicas(out, 1, n_types, counts, masses, box);
}

void md_destroy(md_ctx* c) {
    if (!c) return;
    (void)hipStreamSynchronize(c->st);
    free_all(c);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

int md_get_options(const md_ctx* c, md_options* out) {
    if (!c || !out) return MD_ERR_ARG;
    *out = c->opt;
    return MD_OK;
}

}  // extern "C"

// slots of a group: its tiles of bs records (a block may hold several tiles)
static int64_t group_slots(const md_ctx* c, const GroupInfo& G) { return (G.count + c->bs - 1) / c->bs * c->bs; }

// Record buffer holding each replica's current state.  Replicas may sit at
// different buffers: a split step ends in the buffer it started from while
// an ordinary step ends in the other one.
static int replica_bufs(md_ctx* c, std::vector<int>& cur) {
    std::vector<RepCtrl> C;
    int rc = read_ctrl(c, C);
    if (rc) return rc;
    cur.resize(c->R);
    for (int r = 0; r < c->R; ++r) cur[r] = C[r].cur & 1;
    return MD_OK;
}

// Copy each replica's slot range between the host image and its own buffer.
static int copy_replica_ranges(md_ctx* c, void* host, size_t rsize, const std::vector<int>& cur, bool to_device) {
    for (int r = 0; r < c->R; ++r) {
        const RepInfo& R = c->reps[r];
        if (R.end <= R.start) continue;
        char* h = (char*)host + rsize * (size_t)R.start;
        char* dv = (char*)c->d.rec[cur[r]] + rsize * (size_t)R.start;
        const size_t bytes = rsize * (size_t)(R.end - R.start);
        if (to_device)
            HIPCHK(c, hipMemcpy(dv, h, bytes, hipMemcpyHostToDevice));
        else
            HIPCHK(c, hipMemcpy(h, dv, bytes, hipMemcpyDeviceToHost));
    }
    return MD_OK;
}

// Host <-> device record transfer.  Host arrays are N x 3 (x, y, z per
// particle) in the caller's order; slots add NaN-positioned padding.  With
// `cur` the records go to (and, keeping forces, come from) each replica's
// current buffer; without it to buffer `buf` of every replica.
template <typename FT>
static int put_records(md_ctx* c, int buf, const double* x, const double* v, const double* f,
                       bool keep_forces, const std::vector<int>* cur = nullptr) {
    const size_t N = (size_t)std::max<int64_t>(c->n_slots, 1);
    std::vector<Rec<FT>> h(N);
    if (keep_forces) {
        if (cur) {
            int rc = copy_replica_ranges(c, h.data(), sizeof(Rec<FT>), *cur, false);
            if (rc) return rc;
        } else {
            HIPCHK(c, hipMemcpy(h.data(), c->d.rec[buf], sizeof(Rec<FT>) * N, hipMemcpyDeviceToHost));
        }
    }
    for (const GroupInfo& G : c->groups)
        for (int64_t i = 0; i < group_slots(c, G); ++i) {
            Rec<FT>& r = h[(size_t)(G.start + i)];
            r.type = (uint32_t)G.type;
            if (i >= G.count && !keep_forces) {
                r.x = r.y = r.z = NAN;
                r.vx = r.vy = r.vz = 0.0;
                r.fx = r.fy = r.fz = (FT)0;
            }
        }
    int64_t a = 0;
    for (const GroupInfo& G : c->groups)
        for (int64_t i = 0; i < G.count; ++i, ++a) {
            Rec<FT>& r = h[(size_t)(G.start + i)];
            if (x) {
                r.x = x[3 * a];
                r.y = x[3 * a + 1];
                r.z = x[3 * a + 2];
            }
            if (v) {
                r.vx = v[3 * a];
                r.vy = v[3 * a + 1];
                r.vz = v[3 * a + 2];
            }
            if (f) {
                r.fx = (FT)f[3 * a];
                r.fy = (FT)f[3 * a + 1];
                r.fz = (FT)f[3 * a + 2];
            } else if (!keep_forces) {
                r.fx = r.fy = r.fz = (FT)0;
            }
        }
    if (cur) return copy_replica_ranges(c, h.data(), sizeof(Rec<FT>), *cur, true);
    HIPCHK(c, hipMemcpy(c->d.rec[buf], h.data(), sizeof(Rec<FT>) * N, hipMemcpyHostToDevice));
    return MD_OK;
}

// Records of each replica's current buffer (`cur`, one entry per replica).
template <typename FT>
static int get_records(md_ctx* c, const std::vector<int>& cur, double* x, double* v, double* f) {
    const size_t N = (size_t)std::max<int64_t>(c->n_slots, 1);
    std::vector<Rec<FT>> h(N);
    int rc = copy_replica_ranges(c, h.data(), sizeof(Rec<FT>), cur, false);
    if (rc) return rc;
    int64_t a = 0;
    for (const GroupInfo& G : c->groups)
        for (int64_t i = 0; i < G.count; ++i, ++a) {
            const Rec<FT>& r = h[(size_t)(G.start + i)];
            if (x) {
                x[3 * a] = r.x;
                x[3 * a + 1] = r.y;
                x[3 * a + 2] = r.z;
            }
            if (v) {
                v[3 * a] = r.vx;
                v[3 * a + 1] = r.vy;
                v[3 * a + 2] = r.vz;
            }
            if (f) {
                f[3 * a] = (double)r.fx;
                f[3 * a + 1] = (double)r.fy;
                f[3 * a + 2] = (double)r.fz;
            }
        }
    return MD_OK;
}

static int put_pv(md_ctx* c, const double* pot, const double* vir) {
    const size_t N = (size_t)std::max<int64_t>(c->n_slots, 1);
    std::vector<double> hp(N, 0.0), hv(N, 0.0);
    int64_t a = 0;
    for (const GroupInfo& G : c->groups)
        for (int64_t i = 0; i < G.count; ++i, ++a) {
            if (pot) hp[(size_t)(G.start + i)] = pot[a];
            if (vir) hv[(size_t)(G.start + i)] = vir[a];
        }
    HIPCHK(c, hipMemcpy(c->d.pot, hp.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d.vir, hv.data(), sizeof(double) * N, hipMemcpyHostToDevice));
    return MD_OK;
}

static int get_pv(md_ctx* c, double* pot, double* vir) {
    const size_t N = (size_t)std::max<int64_t>(c->n_slots, 1);
    std::vector<double> hp(N), hv(N);
    HIPCHK(c, hipMemcpy(hp.data(), c->d.pot, sizeof(double) * N, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hv.data(), c->d.vir, sizeof(double) * N, hipMemcpyDeviceToHost));
    int64_t a = 0;
    for (const GroupInfo& G : c->groups)
        for (int64_t i = 0; i < G.count; ++i, ++a) {
            if (pot) pot[a] = hp[(size_t)(G.start + i)];
            if (vir) vir[a] = hv[(size_t)(G.start + i)];
        }
    return MD_OK;
}

// Controls as after an upload of a new state with these boxes (R x 3):
// unit lambda / mu, no lists, input in buffer 0, grid epochs from epoch0.
static int reset_controls(md_ctx* c, const double* box, uint32_t epoch0) {
    std::vector<RepCtrl> C((size_t)c->R);
    for (int r = 0; r < c->R; ++r) {
        RepCtrl X{};
        X.fill_epoch = X.ins_epoch = epoch0;
        for (int k = 0; k < 3; ++k) X.box[k] = box[3 * r + k];
        X.lambda = X.myu = 1.0;
        X.lambda_used = X.myu_used = 1.0;
        X.s = 1.0;
        X.D = 0.0;
        X.flags = 0;
        X.cur = 0;
        C[r] = X;
        for (int k = 0; k < 3; ++k) c->box0[(size_t)r * 3 + k] = box[3 * r + k];
    }
    int rc = write_ctrl(c, C);
    if (rc) return rc;
    c->ctrl_valid = true;
    c->sums_valid = false;
    c->crec_valid = false;
    c->stats.rebuilds = 0;
    c->stats.refresh_steps = 0;
    c->stats.split_steps = 0;
    return MD_OK;
}

// the first grid epoch a reset may use: past every epoch used so far
static int next_epoch(md_ctx* c, uint32_t* epoch0) {
    *epoch0 = 0;
    if (!c->ctrl_valid) return MD_OK;
    std::vector<RepCtrl> old;
    int rc = read_ctrl(c, old);
    if (rc) return rc;
    for (const RepCtrl& X : old) *epoch0 = std::max(*epoch0, std::max(X.fill_epoch, X.ins_epoch));
    *epoch0 += 2u;
    return MD_OK;
}

extern "C" {

int md_upload_state(md_ctx* c, const double* x, const double* v, const double* box) {
    if (!c || !x || !v || !box) return fail(c, MD_ERR_ARG, "NULL argument");
    int rc;
    for (int b = 0; b < 2; ++b) {
        rc = c->fp32 ? put_records<float>(c, b, x, v, nullptr, false)
                     : put_records<double>(c, b, x, v, nullptr, false);
        if (rc) return rc;
    }
    if ((rc = put_pv(c, nullptr, nullptr))) return rc;
    // the grid buffers keep the headers of earlier fills: continue the epoch
    // count past every epoch used so far so that none of them can match again
    uint32_t epoch0 = 0;
    if ((rc = next_epoch(c, &epoch0))) return rc;
    return reset_controls(c, box, epoch0);
}

int md_upload_forces(md_ctx* c, const double* force, const double* pot, const double* vir) {
    if (!c) return MD_ERR_ARG;
    std::vector<int> cur;
    int rc = replica_bufs(c, cur);
    if (rc) return rc;
    std::vector<double> zero;
    if (!force) {
        zero.assign((size_t)c->n_atoms * 3 + 1, 0.0);
        force = zero.data();
    }
    rc = c->fp32 ? put_records<float>(c, 0, nullptr, nullptr, force, true, &cur)
                 : put_records<double>(c, 0, nullptr, nullptr, force, true, &cur);
    if (rc) return rc;
    if ((rc = put_pv(c, pot, vir))) return rc;
    c->sums_valid = false;
    c->crec_valid = false;
    return MD_OK;
}

int md_download_state(md_ctx* c, double* x, double* v, double* force, double* pot, double* vir,
                      double* box) {
    if (!c) return MD_ERR_ARG;
    std::vector<RepCtrl> C;
    int rc = read_ctrl(c, C);
    if (rc) return rc;
    if (x || v || force) {
        std::vector<int> cur(c->R);
        for (int r = 0; r < c->R; ++r) cur[r] = C[r].cur & 1;
        rc = c->fp32 ? get_records<float>(c, cur, x, v, force) : get_records<double>(c, cur, x, v, force);
        if (rc) return rc;
    }
    if (pot || vir)
        if ((rc = get_pv(c, pot, vir))) return rc;
    if (box)
        for (int r = 0; r < c->R; ++r)
            for (int k = 0; k < 3; ++k) box[3 * r + k] = C[r].box[k];
    return MD_OK;
}

// Relayout keeping the state (used when block size / force path change).
static int relayout(md_ctx* c, const md_options& nopt) {
    const int64_t n = c->n_atoms;
    std::vector<double> x((size_t)n * 3 + 1), v((size_t)n * 3 + 1), f((size_t)n * 3 + 1),
        pot((size_t)n + 1), vir((size_t)n + 1), box((size_t)c->R * 3);
    int rc = md_download_state(c, x.data(), v.data(), f.data(), pot.data(), vir.data(), box.data());
    if (rc) return rc;
    std::vector<RepCtrl> C;
    if ((rc = read_ctrl(c, C))) return rc;
    c->opt = nopt;
    c->bs = nopt.block_size;
    for (int r = 0; r < c->R; ++r)
        for (int k = 0; k < 3; ++k) c->box0[(size_t)r * 3 + k] = box[(size_t)r * 3 + k];
    if ((rc = alloc_base(c))) return rc;
    if ((rc = md_upload_state(c, x.data(), v.data(), box.data()))) return rc;
    // restore forces, potentials and virials in the new precision
    if ((rc = md_upload_forces(c, f.data(), pot.data(), vir.data()))) return rc;
    std::vector<RepCtrl> C2;
    if ((rc = read_ctrl(c, C2))) return rc;
    for (int r = 0; r < c->R; ++r) {
        C2[r].lambda = C[r].lambda;
        C2[r].myu = C[r].myu;
        C2[r].psi = C[r].psi;
        C2[r].lambda_used = C[r].lambda_used;
        C2[r].myu_used = C[r].myu_used;
        C2[r].psi_used = C[r].psi_used;
        C2[r].step = C[r].step;
        C2[r].rebuilds = C[r].rebuilds;
        C2[r].nulls = C[r].nulls;
    }
    return write_ctrl(c, C2);
}

int md_set_options(md_ctx* c, const md_options* o) {
    if (!c || !o) return fail(c, MD_ERR_ARG, "NULL argument");
    md_options n = *o;
    if (n.force_mode < 0 || n.force_mode > 3) return fail(c, MD_ERR_ARG, "bad force_mode");
    if (n.force_mode == MD_FORCE_MLP && c->T != 1)
        return fail(c, MD_ERR_UNIMPLEMENTED, "the surrogate is defined for one particle type");
    // the surrogate's pair and feature kernels run fastest at 512 threads
    // (262,144-atom liquid: 0.301 -> 0.287 ms per step); the LJ paths at 256
    if (n.block_size == 0) n.block_size = n.force_mode == MD_FORCE_MLP ? 512 : 256;
    if (n.block_size != 64 && n.block_size != 128 && n.block_size != 256 && n.block_size != 512)
        return fail(c, MD_ERR_ARG, "block_size must be 64, 128, 256 or 512");
    if (!(n.cells_per_atom > 0)) n.cells_per_atom = 2.0;
    if (n.graph_chunk < 0) return fail(c, MD_ERR_ARG, "graph_chunk < 0");
    if (n.skin < 0) return fail(c, MD_ERR_ARG, "skin < 0");
    if (n.tiles_per_block < 0 || n.tiles_per_block > 64)
        return fail(c, MD_ERR_ARG, "tiles_per_block must be 0 (automatic) or 1..64");
    c->skin_set = n.skin > 0;
    c->crec_valid = false;
    const bool relay = n.block_size != c->opt.block_size || n.force_mode != c->opt.force_mode ||
                       n.cells_per_atom != c->opt.cells_per_atom ||
                       n.tiles_per_block != c->opt.tiles_per_block;
    if (relay) {
        int rc = relayout(c, n);
        if (rc) return rc;
    } else {
        c->opt = n;
    }
    rebuild_pairs(c);
    if (n.list_capacity > 0) c->d.list_cap = n.list_capacity;
    if (n.cell_capacity > 0) c->d.cell_cap = n.cell_capacity;
    if (n.list_capacity <= 0 || n.cell_capacity <= 0) auto_capacities(c);
    int rc = alloc_lists(c);
    if (rc) return rc;
    rc = upload_pairs(c);
    if (rc) return rc;
    std::vector<RepCtrl> C;
    if ((rc = read_ctrl(c, C))) return rc;
    for (auto& X : C) X.flags &= ~(F_HAVE_LIST | F_GRID_FRESH);
    c->opt.skin = c->opt.skin;  // resolved by rebuild_pairs when unset
    return write_ctrl(c, C);
}

static int set_pot(md_ctx* c, uint16_t t0, uint16_t t1) {
    (void)t0;
    (void)t1;
    c->crec_valid = false;
    rebuild_pairs(c);
    int rc = upload_pairs(c);
    if (rc) return rc;
    std::vector<RepCtrl> C;
    if ((rc = read_ctrl(c, C))) return rc;
    for (auto& X : C) X.flags &= ~(F_HAVE_LIST | F_GRID_FRESH);
    c->sums_valid = false;
    return write_ctrl(c, C);
}

int md_set_pair_potential(md_ctx* c, uint16_t t0, uint16_t t1, double sigma, double eps) {
    if (!c || t0 >= c->T || t1 >= c->T) return fail(c, MD_ERR_ARG, "bad type id");
    const std::pair<int, int> key(std::min(t0, t1), std::max(t0, t1));
    c->pot_lj[key] = {sigma, eps};
    c->pot_cut.erase(key);
    return set_pot(c, t0, t1);
}

int md_set_pair_potential_lj(md_ctx* c, uint16_t t0, uint16_t t1, double sigma, double eps,
                             double r_cut, double u_cut) {
    if (!c || t0 >= c->T || t1 >= c->T) return fail(c, MD_ERR_ARG, "bad type id");
    if (!(r_cut > 0)) return fail(c, MD_ERR_ARG, "r_cut must be positive");
    const std::pair<int, int> key(std::min(t0, t1), std::max(t0, t1));
    c->pot_lj[key] = {sigma, eps};
    c->pot_cut[key] = {r_cut, u_cut};
    return set_pot(c, t0, t1);
}

int md_update_force(md_ctx* c) {
    if (!c) return MD_ERR_ARG;
    c->crec_valid = false;
    if (c->mlp && !c->mlp_ready) return fail(c, MD_ERR_ARG, "surrogate parameters not set (md_mlp_set_params)");
    if (c->list_mode) return list_force_eval(c);
    std::vector<RepParams> P((size_t)c->R, base_params(c));
    int rc = write_params(c, P);
    if (rc) return rc;
    c->d.par = c->par;
    if (c->n_blocks > 0) HIPCHK(c, launch_allpairs(c->d, c->bs, true, c->st));
    HIPCHK(c, finalize(c, M_REDUCE));
    if ((rc = reduce_state_sums(c, 1))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    c->sums_valid = true;
    return MD_OK;
}

int md_apply_boundary_conditions(md_ctx* c) {
    if (!c) return MD_ERR_ARG;
    c->crec_valid = false;
    c->d.par = c->par;
    if (c->n_blocks > 0) HIPCHK(c, launch_wrap(c->d, c->fp32, c->bs, c->st));
    HIPCHK(c, hipStreamSynchronize(c->st));
    return MD_OK;
}

int md_step(md_ctx* c, double dt, md_thermostat* th, md_barostat* ba, int64_t n_steps) {
    if (!c) return MD_ERR_ARG;
    return run_steps(c, dt, th, ba, n_steps, false);
}

int md_step_replicas(md_ctx* c, double dt, md_thermostat* th, md_barostat* ba, int64_t n_steps) {
    if (!c) return MD_ERR_ARG;
    return run_steps(c, dt, th, ba, n_steps, true);
}

// Temperature and pressure of (replica 0, type) from the device sums.
static int type_thermo(md_ctx* c, int rep, int type, double* temperature, double* pressure,
                       double out[11]) {
    int rc = ensure_sums(c);
    if (rc) return rc;
    std::vector<double> S;
    if ((rc = read_sums(c, S))) return rc;
    std::vector<RepCtrl> C;
    if ((rc = read_ctrl(c, C))) return rc;
    const int g = c->reps[rep].first_group + type;
    const GroupInfo& G = c->groups[g];
    const double* s = &S[(size_t)g * NS];
    const double M = G.sum_m;
    const double cx = s[S_MVX] / M, cy = s[S_MVY] / M, cz = s[S_MVZ] / M;
    // the stored sums hold sum m |v - v_com|^2 (energy.rs:25-37)
    const double th2 = s[S_MV2];
    const double thermal = th2 / 2.0;
    const double t = (2.0 * thermal) / (3.0 * (double)G.count * MD_K_B) * 100.0;
    const double* b = C[rep].box;
    const double vol = b[0] * b[1] * b[2];
    const double p = (th2 + (-s[S_VIR]) * 0.5) / vol / 3.0;
    if (temperature) *temperature = t;
    if (pressure) *pressure = p;
    if (out) {
        // energy.rs:14-22: sum m |v|^2 = sum m |v - v_com|^2 + M |v_com|^2
        out[0] = (th2 + M * ((cx * cx + cy * cy) + cz * cz)) / 2.0;
        out[1] = s[S_POT] / 2.0;
        out[2] = thermal;
        out[3] = t;
        out[4] = p;
        out[5] = s[S_MVX];
        out[6] = s[S_MVY];
        out[7] = s[S_MVZ];
        out[8] = cx;
        out[9] = cy;
        out[10] = cz;
    }
    return MD_OK;
}

int md_macro(md_ctx* c, int replica, uint16_t type_id, double out[11]) {
    if (!c || !out || replica < 0 || replica >= c->R || type_id >= c->T)
        return fail(c, MD_ERR_ARG, "bad argument");
    return type_thermo(c, replica, type_id, nullptr, nullptr, out);
}

int md_thermostat_calculate_lambda(md_ctx* c, md_thermostat* th, double dt, uint16_t type_id) {
    if (!c || !th || type_id >= c->T) return fail(c, MD_ERR_ARG, "bad argument");
    if (th->kind == 3) return fail(c, MD_ERR_UNIMPLEMENTED, "Thermostat::Custom is todo!() in the reference");
    double t;
    int rc = type_thermo(c, 0, type_id, &t, nullptr, nullptr);
    if (rc) return rc;
    if (th->kind == 1) {
        th->lambda = sqrt(1.0 + dt / th->tau * (th->target_t / t - 1.0));
    } else if (th->kind == 2) {
        const double psi_dot = -((th->target_t / t) - 1.0) / th->tau;
        th->psi += psi_dot * (dt / 2.0);
        th->lambda = exp(-th->psi * dt / 2.0);
    }
    return MD_OK;
}

int md_thermostat_update(md_ctx* c, md_thermostat* th, double dt, uint16_t type_id) {
    if (!c || !th || type_id >= c->T) return fail(c, MD_ERR_ARG, "bad argument");
    if (th->kind == 3) return fail(c, MD_ERR_UNIMPLEMENTED, "Thermostat::Custom is todo!() in the reference");
    double t;
    int rc = type_thermo(c, 0, type_id, &t, nullptr, nullptr);
    if (rc) return rc;
    if (th->kind == 1 || th->kind == 2) {
        c->crec_valid = false;
        const GroupInfo& G = c->groups[c->reps[0].first_group + type_id];
        c->d.par = c->par;
        HIPCHK(c, launch_scale(c->d, c->fp32, G.start, G.count, th->lambda, 0, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
        c->sums_valid = false;
    }
    if (th->kind == 2) {
        const double psi_dot = -((th->target_t / t) - 1.0) / th->tau;
        th->psi += psi_dot * (dt / 2.0);
    }
    return MD_OK;
}

int md_barostat_calculate_myu(md_ctx* c, md_barostat* ba, double dt, uint16_t type_id) {
    if (!c || !ba || type_id >= c->T) return fail(c, MD_ERR_ARG, "bad argument");
    if (ba->kind == 2) return fail(c, MD_ERR_UNIMPLEMENTED, "Barostat::Custom is todo!() in the reference");
    double p;
    int rc = type_thermo(c, 0, type_id, nullptr, &p, nullptr);
    if (rc) return rc;
    if (ba->kind == 1) ba->myu = cbrt(1.0 + dt * ba->beta / ba->tau * (p - ba->target_p));
    return MD_OK;
}

int md_barostat_update(md_ctx* c, md_barostat* ba, uint16_t type_id) {
    if (!c || !ba || type_id >= c->T) return fail(c, MD_ERR_ARG, "bad argument");
    if (ba->kind == 2) return fail(c, MD_ERR_UNIMPLEMENTED, "Barostat::Custom is todo!() in the reference");
    if (ba->kind != 1) return MD_OK;
    c->crec_valid = false;
    const GroupInfo& G = c->groups[c->reps[0].first_group + type_id];
    c->d.par = c->par;
    HIPCHK(c, launch_scale(c->d, c->fp32, G.start, G.count, ba->myu, 1, c->st));
    std::vector<RepCtrl> C;
    int rc = read_ctrl(c, C);
    if (rc) return rc;
    for (int k = 0; k < 3; ++k) C[0].box[k] *= ba->myu;
    C[0].flags &= ~(F_HAVE_LIST | F_GRID_FRESH);
    c->sums_valid = false;
    return write_ctrl(c, C);
}

int md_init_lattice(int kind, const double start[3], const int64_t grid[3], double a, double* pos) {
    if (!start || !grid || !pos || (kind != 0 && kind != 1)) return fail(nullptr, MD_ERR_ARG, "bad argument");
    // initializer/position.rs:34-99: index x * ny * nz + y * nz + z
    for (int64_t x = 0; x < grid[0]; ++x)
        for (int64_t y = 0; y < grid[1]; ++y)
            for (int64_t z = 0; z < grid[2]; ++z) {
                const int64_t cell = x * grid[1] * grid[2] + y * grid[2] + z;
                const double X = (double)x, Y = (double)y, Z = (double)z;
                if (kind == 0) {
                    double* p = pos + 3 * cell;
                    p[0] = start[0] + X * a;
                    p[1] = start[1] + Y * a;
                    p[2] = start[2] + Z * a;
                } else {
                    double* p = pos + 12 * cell;
                    const double sites[4][3] = {{X, Y, Z}, {X, Y + 0.5, Z + 0.5}, {X + 0.5, Y, Z + 0.5},
                                                {X + 0.5, Y + 0.5, Z}};
                    for (int k = 0; k < 4; ++k)
                        for (int dd = 0; dd < 3; ++dd) p[3 * k + dd] = start[dd] + sites[k][dd] * a;
                }
            }
    return MD_OK;
}

// SplitMix64 stream and Box-Muller pairs (cos draw first, then sin).
static uint64_t sm64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double open01(uint64_t* s) {
    return ((double)(sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

int md_init_mb_velocities(double* vel, int64_t n, double temperature, double mass, uint64_t seed) {
    if (!vel || n < 0 || !(mass > 0)) return fail(nullptr, MD_ERR_ARG, "bad argument");
    // initializer/velocity.rs:6-29: sigma = sqrt(K_B * T * 0.01 / m), i and i + n/2 opposite
    const double tp = temperature * 0.01;
    const double sigma = sqrt(MD_K_B * tp / mass);
    uint64_t s = seed;
    bool have = false;
    double spare = 0.0;
    auto normal = [&]() {
        if (have) {
            have = false;
            return spare;
        }
        const double u1 = open01(&s), u2 = open01(&s);
        const double r = sqrt(-2.0 * log(u1));
        const double th = 6.283185307179586 * u2;
        spare = r * sin(th);
        have = true;
        return r * cos(th);
    };
    memset(vel, 0, sizeof(double) * 3 * (size_t)n);
    const int64_t h = n / 2;
    for (int64_t i = 0; i < h; ++i) {
        const double vx = sigma * normal();
        const double vy = sigma * normal();
        const double vz = sigma * normal();
        vel[3 * i] = vx;
        vel[3 * i + 1] = vy;
        vel[3 * i + 2] = vz;
        vel[3 * (i + h)] = -vx;
        vel[3 * (i + h) + 1] = -vy;
        vel[3 * (i + h) + 2] = -vz;
    }
    return MD_OK;
}

// ------------------------------------------------------------------ device initialisers

// The records to edit on the device: each replica's current buffer and the
// boxes; a context that never held a state gets the blank one first
// (initialize_particles).
static int edit_begin(md_ctx* c, std::vector<int>& cur, std::vector<double>& boxes) {
    boxes.assign((size_t)c->R * 3, 0.0);
    cur.assign(c->R, 0);
    if (!c->ctrl_valid) {
        HIPCHK(c, launch_init_blank(c->d, c->fp32, c->d.groups, (int)c->groups.size(), c->n_slots, c->st));
        HIPCHK(c, hipMemsetAsync(c->d.pot, 0, sizeof(double) * (size_t)std::max<int64_t>(c->n_slots, 1), c->st));
        HIPCHK(c, hipMemsetAsync(c->d.vir, 0, sizeof(double) * (size_t)std::max<int64_t>(c->n_slots, 1), c->st));
        for (size_t k = 0; k < boxes.size(); ++k) boxes[k] = c->box0[k];
        return reset_controls(c, boxes.data(), 0);
    }
    std::vector<RepCtrl> C;
    int rc = read_ctrl(c, C);
    if (rc) return rc;
    for (int r = 0; r < c->R; ++r) {
        cur[r] = C[(size_t)r].cur & 1;
        for (int k = 0; k < 3; ++k) boxes[(size_t)r * 3 + k] = C[(size_t)r].box[k];
    }
    return MD_OK;
}

// After an edit of each replica's buffer `cur[r]`: both buffers hold the
// edited state, per-atom potentials and virials cleared, controls reset as
// by md_upload_state.
static int edit_commit(md_ctx* c, const std::vector<int>& cur, const std::vector<double>& boxes) {
    const size_t rsize = c->fp32 ? sizeof(Rec<float>) : sizeof(Rec<double>);
    for (int r = 0; r < c->R; ++r) {
        const RepInfo& R = c->reps[r];
        if (R.end <= R.start) continue;
        const size_t off = rsize * (size_t)R.start, bytes = rsize * (size_t)(R.end - R.start);
        char* src = (char*)c->d.rec[cur[r]] + off;
        char* dst = (char*)c->d.rec[cur[r] ^ 1] + off;
        HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->st));
    }
    HIPCHK(c, hipMemsetAsync(c->d.pot, 0, sizeof(double) * (size_t)std::max<int64_t>(c->n_slots, 1), c->st));
    HIPCHK(c, hipMemsetAsync(c->d.vir, 0, sizeof(double) * (size_t)std::max<int64_t>(c->n_slots, 1), c->st));
    uint32_t epoch0 = 0;
    int rc = next_epoch(c, &epoch0);
    if (rc) return rc;
    return reset_controls(c, boxes.data(), epoch0);
}

static const GroupInfo* find_group(md_ctx* c, int replica, int type_id) {
    for (const GroupInfo& G : c->groups)
        if (G.rep == replica && G.type == type_id) return &G;
    return nullptr;
}

int md_init_positions(md_ctx* c, int replica, uint16_t type_id, int kind, const double start[3],
                      const int64_t grid[3], double a) {
    if (!c || !start || !grid || (kind != 0 && kind != 1)) return fail(c, MD_ERR_ARG, "bad argument");
    const GroupInfo* G = find_group(c, replica, (int)type_id);
    if (!G) return fail(c, MD_ERR_ARG, "InitError::ParticleIdDidNotFound: no such replica or type");
    if (grid[0] < 0 || grid[1] < 0 || grid[2] < 0) return fail(c, MD_ERR_ARG, "negative grid size");
    std::vector<int> cur;
    std::vector<double> boxes;
    int rc = edit_begin(c, cur, boxes);
    if (rc) return rc;
    const double* box = &boxes[(size_t)replica * 3];
    // position.rs:166-171, then :35-36 / :53-54
    if ((double)grid[0] * a > box[0] || (double)grid[1] * a > box[1] || (double)grid[2] * a > box[2])
        return fail(c, MD_ERR_ARG, "InitError::OutOfBoundary: the lattice is larger than the box");
    const int64_t sites = grid[0] * grid[1] * grid[2] * (kind == 1 ? 4 : 1);
    if (sites > G->count) return fail(c, MD_ERR_ARG, "InitError::TooBig: more lattice sites than particles");
    HIPCHK(c, launch_init_lattice(c->d.rec[cur[replica]], c->fp32, G->start, sites, kind == 1, start, grid, a, c->st));
    return edit_commit(c, cur, boxes);
}

int md_init_velocities(md_ctx* c, int replica, uint16_t type_id, double temperature, uint64_t seed) {
    if (!c) return MD_ERR_ARG;
    const GroupInfo* G = find_group(c, replica, (int)type_id);
    if (!G) return fail(c, MD_ERR_ARG, "no such replica or type");
    std::vector<int> cur;
    std::vector<double> boxes;
    int rc = edit_begin(c, cur, boxes);
    if (rc) return rc;
    // velocity.rs:8-10
    const double tp = temperature * 0.01;
    const double sigma = sqrt(MD_K_B * tp / G->mass);
    HIPCHK(c, launch_init_mb(c->d.rec[cur[replica]], c->fp32, G->start, G->count, sigma, seed, c->st));
    return edit_commit(c, cur, boxes);
}

int md_get_stats(md_ctx* c, md_stats* out) {
    if (!c || !out) return MD_ERR_ARG;
    md_stats s = c->stats;
    s.steps = c->stats.steps;
    s.rebuilds = c->last_rebuilds;
    s.refresh_steps = c->last_refresh;
    s.split_steps = c->last_splits;
    s.list_capacity = c->d.list_cap;
    s.cell_capacity = c->d.cell_cap;
    s.n_slots = (int32_t)c->n_slots;
    s.mean_list_len = 0.0;
    s.max_list_len = 0;
    s.list_visits = c->list_visits;
    if (c->list_mode && c->n_slots > 0) {
        std::vector<int32_t> nc((size_t)c->n_slots);
        if (hipMemcpy(nc.data(), c->d.ncount, sizeof(int32_t) * nc.size(), hipMemcpyDeviceToHost) == hipSuccess) {
            double sum = 0;
            int mx = 0;
            for (const GroupInfo& G : c->groups)
                for (int64_t i = 0; i < G.count; ++i) {
                    const int v = nc[(size_t)(G.start + i)];
                    sum += v;
                    mx = std::max(mx, v);
                }
            s.mean_list_len = c->n_atoms ? sum / (double)c->n_atoms : 0.0;
            s.max_list_len = mx;
        }
    }
    *out = s;
    return MD_OK;
}

int md_mlp_set_params(md_ctx* c, const float* theta, const float* g_mean, const float* g_inv_std,
                      const float* g_lo, const float* g_hi, double e_scale, double e_shift,
                      int judge_every) {
    if (!c || !theta || !g_mean || !g_inv_std) return fail(c, MD_ERR_ARG, "NULL argument");
    if (judge_every < 0) return fail(c, MD_ERR_ARG, "judge_every < 0");
    c->mlp_theta.assign(theta, theta + kMlpNP);
    MlpConst& k = c->mlp_const;
    k = MlpConst{};
    // radial basis of the Lennard-Jones sigma of type pair (0, 0)
    const double sigma = c->pairs[0].sigma;
    const double w = 0.1 * sigma;
    for (int q = 0; q < kMlpK; ++q) {
        k.g_mean[q] = g_mean[q];
        k.g_inv_std[q] = g_inv_std[q];
        k.g_lo[q] = g_lo ? g_lo[q] : -INFINITY;
        k.g_hi[q] = g_hi ? g_hi[q] : INFINITY;
        k.mu[q] = (float)(sigma * (0.8 + 1.7 * q / (kMlpK - 1)));
    }
    k.inv_2w2 = (float)(1.0 / (2.0 * w * w));
    k.inv_w2 = (float)(1.0 / (w * w));
    {
        const double a = 1.0 / (2.0 * w * w), dm = sigma * 1.7 / (kMlpK - 1);
        k.bas_2ad = (float)(2.0 * a * dm);
        k.bas_ad2 = (float)(a * dm * dm);
        k.bas_c = (float)exp(-2.0 * a * dm * dm);
    }
    k.rc = (float)(2.5 * sigma);
    k.pi_rc = (float)(M_PI / (2.5 * sigma));
    k.e_scale = (float)e_scale;
    k.e_shift = (float)e_shift;
    k.b3 = theta[kMlpOffB3];
    k.judge_every = judge_every;
    k.check_range = (g_lo && g_hi) ? 1 : 0;
    if ((double)k.rc > c->rc_max)
        return fail(c, MD_ERR_ARG, "surrogate cutoff exceeds the neighbour-list cutoff");
    c->mlp_ready = true;
    c->sums_valid = false;
    int rc = mlp_upload(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->st));
    return MD_OK;
}

int md_mlp_get_stats(md_ctx* c, int replica, double out[5]) {
    if (!c || !out || replica < 0 || replica >= c->R) return fail(c, MD_ERR_ARG, "bad argument");
    if (!c->mlp || !c->d.mlp_stat) return fail(c, MD_ERR_ARG, "force mode is not MD_FORCE_MLP");
    MlpStat m;
    HIPCHK(c, hipMemcpy(&m, c->d.mlp_stat + replica, sizeof(MlpStat), hipMemcpyDeviceToHost));
    const double n = (double)std::max<int64_t>(c->reps[replica].n_atoms, 1);
    out[0] = m.last_ref2 > 0.0 ? sqrt(m.last_err2 / m.last_ref2) : 0.0;
    out[1] = m.max_rel;
    out[2] = (double)m.checks;
    out[3] = (double)m.fallbacks;
    out[4] = sqrt(m.last_ref2 / (3.0 * n));
    return MD_OK;
}

int md_set_profiling(md_ctx* c, int on) {
    if (!c) return MD_ERR_ARG;
    for (int k = 0; k < P_NUM; ++k) {
        c->ktime[k] = c->kspan_time[k] = 0.0;
        c->kcount[k] = c->kspan_count[k] = 0;
    }
    c->list_visits = 0;
    if (on && !c->prof_ring) {
        void* p = nullptr;
        HIPCHK(c, hipMalloc(&p, sizeof(unsigned long long) * ((size_t)kProfCount + 1)));
        c->prof_ring = (unsigned long long*)p;
        HIPCHK(c, hipMalloc(&p, sizeof(unsigned long long) * 3 * P_NUM * (size_t)std::max(c->n_blocks, 1)));
        c->prof_blk = (unsigned long long*)p;
    }
    c->prof = on != 0;
    c->d.prof = c->prof ? c->prof_ring : nullptr;
    c->d.prof_blk = c->prof ? c->prof_blk : nullptr;
    drop_graph(c);  // the graphs hold the old kernel arguments
    if (c->prof) {
        HIPCHK(c, launch_prof_init(c->d, c->st));
        HIPCHK(c, hipStreamSynchronize(c->st));
    }
    return MD_OK;
}

int md_kernel_time(md_ctx* c, int kernel, double* total_ms, int64_t* launches, double* span_ms,
                   int64_t* span_launches) {
    if (!c || kernel < 0 || kernel >= K_NUM) return MD_ERR_ARG;
    // the step slot holds k_step on the list paths and k_allpairs on the plain path
    int slot = -1;
    if (kernel == K_STEP && c->list_mode) slot = P_STEP;
    if (kernel == K_ALLPAIRS && !c->list_mode) slot = P_STEP;
    if (kernel == K_FINALIZE) slot = P_FIN;
    if (kernel == K_DRIFT && (!c->list_mode || c->mlp)) slot = P_DRIFT;
    if (kernel == K_MLP_NET && c->mlp) slot = P_NET;
    if (kernel == K_MLP_FEAT && c->mlp) slot = P_FEAT;
    if (total_ms) *total_ms = slot >= 0 ? c->ktime[slot] : 0.0;
    if (launches) *launches = slot >= 0 ? c->kcount[slot] : 0;
    if (span_ms) *span_ms = slot >= 0 ? c->kspan_time[slot] : 0.0;
    if (span_launches) *span_launches = slot >= 0 ? c->kspan_count[slot] : 0;
    return MD_OK;
}

}  // extern "C"
