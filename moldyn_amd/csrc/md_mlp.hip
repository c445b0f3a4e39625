// md_mlp.hip -- learned per-atom energy surrogate on gfx950 (inference).
//
// E_i = e_scale * MLP(x_i) + e_shift with x_i = (G_i - mean) * inv_std, where
// G_ik = sum_j exp(-(r_ij - mu_k)^2 / (2 w^2)) f_c(r_ij) (32 radial features,
// cosine cutoff f_c) and MLP = 32-128-128-1 with SiLU.  Forces are the exact
// negative gradient of sum_i E_i:
//   F_i = sum_j (c_i(r_ij) + c_j(r_ij)) d_ij / r_ij,  c_i(r) = sum_k dE_i/dG_ik dg_k/dr.
// A surrogate step launch is five kernels:
//   k_mlp_desc   (PHASE 1) half kick, thermostat scale, drift, wrap
//                (integrator.rs:28-45) into buffer cur ^ 1, and FP32 copies
//                of the drifted positions;
//   k_mlp_desc   (PHASE 2) the features of the drifted positions, and the
//                list entries within the cutoff;
//   k_mlp_net    per 16-atom row tile of a wave: the forward pass and the
//                backward pass to dE/dG on BF16 MFMA (FP32 accumulation),
//                transposed so that activations stay in registers, weights
//                staged in LDS once per block, one 16-wave block per CU;
//   k_mlp_force  pair forces from both atoms' dE/dG, the LJ fallback when a
//                feature left the training range, the periodic LJ check,
//                second half kick and barostat scale into buffer cur;
//   k_fin        the usual finalize.
#include "md_common.h"

namespace md {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float silu_f(float a, float& ds) {
    // the hardware reciprocal (1 ulp): the activations are rounded to BF16
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-a));
    ds = s * (1.0f + a * (1.0f - s));
    return a * s;
}

// The 32 Gaussians exp(-(r - mu_k)^2 / (2 w^2)) of one distance, handed
// to f(k, value) in order of k.  The centres are evenly spaced, so within
// each block of 16 the values follow from the block's first one by the
// recurrence g_{k+1} = g_k q_k, q_{k+1} = q_k exp(-2 a D^2) (MlpConst): four
// exponentials instead of 32, and no array of values.  A block's first value
// is at least exp(-a (15 D)^2) ~ 1e-15 whenever the distance lies among its
// centres, so nothing that matters underflows.
template <typename F>
__device__ __forceinline__ void gauss_basis(float r, const MlpConst& mc, F&& f) {
#pragma unroll
    for (int k0 = 0; k0 < kMlpK; k0 += 16) {
        const float t = r - mc.mu[k0];
        float g = __expf(-t * t * mc.inv_2w2);
        float q = __expf(fmaf(mc.bas_2ad, t, -mc.bas_ad2));
#pragma unroll
        for (int k = k0; k < k0 + 16; ++k) {
            f(k, g);
            g *= q;
            q *= mc.bas_c;
        }
    }
}

// Features of one particle from its neighbour list and FP64 records (the
// force-only evaluation, at the current positions).
__device__ __forceinline__ void mlp_features(const DevState& d, const MlpConst& mc, const Rec<float>* __restrict__ pos,
                                             int64_t s, double x1, double y1, double z1, double bx, double by,
                                             double bz, float (&g)[kMlpK]) {
#pragma unroll
    for (int k = 0; k < kMlpK; ++k) g[k] = 0.0f;
    const int nn = d.ncount[s];
    const double rc2 = (double)mc.rc * (double)mc.rc;
    const int64_t ns = d.n_slots;
    // software pipeline over the list: the index two entries ahead, the
    // neighbour's record one entry ahead
    int32_t jn = nn > 1 ? d.nbr[ns + s] : 0;
    Rec<float> on;
    if (nn > 0) on = pos[d.nbr[s]];
    for (int q = 0; q < nn; ++q) {
        const Rec<float> o = on;
        if (q + 1 < nn) {
            on = pos[jn];
            jn = q + 2 < nn ? d.nbr[(int64_t)(q + 2) * ns + s] : 0;
        }
        const double dx = min_image(o.x - x1, bx);
        const double dy = min_image(o.y - y1, by);
        const double dz = min_image(o.z - z1, bz);
        const double r2 = (dx * dx + dy * dy) + dz * dz;
        if (!(r2 < rc2)) continue;
        // FP32 distance and cutoff from the hardware square root (1 ulp)
        const float r = __builtin_amdgcn_sqrtf((float)r2);
        const float fc = 0.5f * (__cosf(r * mc.pi_rc) + 1.0f);
        gauss_basis(r, mc, [&](int k, float e) { g[k] = fmaf(e, fc, g[k]); });
    }
}

// The same from the FP32 copies of the drifted positions (mlp_pos) that a
// stepping launch writes before its features: 16 bytes per neighbour
// instead of 24 and four particles per 64-byte line instead of one.  FP32
// coordinates (absolute error below 5e-7 nm in a 16 nm box) perturb the
// distances by less than 1e-5 relative, well under the BF16 rounding of the
// features and of dE/dG that follows.  The entries within r_c, in list
// order, go to mlp_nbr / mlp_ncount for the force kernel.
__device__ __forceinline__ void mlp_features_f32(const DevState& d, const MlpConst& mc, int64_t s, float bx, float by,
                                                 float bz, float (&g)[kMlpK]) {
#pragma unroll
    for (int k = 0; k < kMlpK; ++k) g[k] = 0.0f;
    const float4* __restrict__ pos = d.mlp_pos;
    const float4 a = pos[s];
    const int nn = d.ncount[s];
    const float rc2 = mc.rc * mc.rc;
    const int64_t ns = d.n_slots;
    int32_t jn = nn > 1 ? d.nbr[ns + s] : 0;  // index of the entry after the next
    int32_t j1 = 0;                           // index of the next entry (its position is in on)
    float4 on;
    if (nn > 0) {
        j1 = d.nbr[s];
        on = pos[j1];
    }
    int kept = 0;
    for (int q = 0; q < nn; ++q) {
        const float4 o = on;
        const int32_t j = j1;
        if (q + 1 < nn) {
            j1 = jn;
            on = pos[jn];
            jn = q + 2 < nn ? d.nbr[(int64_t)(q + 2) * ns + s] : 0;
        }
        const float dx = min_image_f(o.x - a.x, bx);
        const float dy = min_image_f(o.y - a.y, by);
        const float dz = min_image_f(o.z - a.z, bz);
        const float r2 = (dx * dx + dy * dy) + dz * dz;
        if (!(r2 < rc2)) continue;
        d.mlp_nbr[(int64_t)kept++ * ns + s] = j;
        const float r = __builtin_amdgcn_sqrtf(r2);
        const float fc = 0.5f * (__cosf(r * mc.pi_rc) + 1.0f);
        gauss_basis(r, mc, [&](int k, float e) { g[k] = fmaf(e, fc, g[k]); });
    }
    d.mlp_ncount[s] = kept;
}

__device__ __forceinline__ bool mlp_store_features(const DevState& d, const MlpConst& mc, int64_t s,
                                                   const float (&g)[kMlpK], bool check) {
    bool oor = false;
    bf16x8 v[kMlpK / 8];
#pragma unroll
    for (int k = 0; k < kMlpK; ++k) {
        if (check && mc.check_range && (g[k] < mc.g_lo[k] || g[k] > mc.g_hi[k])) oor = true;
        v[k / 8][k % 8] = (__bf16)((g[k] - mc.g_mean[k]) * mc.g_inv_std[k]);
    }
    bf16x8* out = (bf16x8*)(d.mlp_x + s * kMlpK);
#pragma unroll
    for (int q = 0; q < kMlpK / 8; ++q) out[q] = v[q];
    return oor;
}

// Features of every particle, and in a stepping launch the first half of
// the velocity-Verlet step.  PHASE 0 (FORCE_ONLY) computes the features at
// the current positions from the FP64 records.  A stepping launch runs
// PHASE 1, the first half step into records cur ^ 1 and the FP32 position
// copies mlp_pos (a split step's second launch, which does not move the
// particles, only copies its drifted positions there), and then PHASE 2, the
// features from mlp_pos.
template <int BS, bool FORCE_ONLY, int PHASE>
__global__ __launch_bounds__(BS) void k_mlp_desc(DevState d) {
    static_assert(FORCE_ONLY == (PHASE == 0), "PHASE 0 is the force-only evaluation");
    constexpr bool kDrift = PHASE == 1, kFeat = PHASE != 1;
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const int r = G.rep;
    const RepCtrl& C = ctrl_in(d, r);
    const uint32_t flags = FORCE_ONLY ? 0u : C.flags;
    if (flags & (F_EXIT | F_BUILD)) {
        if (PHASE != 2) prof_block(d, P_DRIFT, 0, false, 0);
        return;
    }
    const uint64_t t_start = prof_now(d);
    const MlpConst& mc = *d.mlp_c;
    const RepParams& P = d.params[r];
    const RepInfo R = d.reps[r];
    const TileRange tr = block_tiles(d, G, b, BS);
    const int lane = threadIdx.x;
    const int cur = C.cur, nx = cur ^ 1;
    const Rec<float>* __restrict__ rin = (const Rec<float>*)d.rec[cur];
    Rec<float>* __restrict__ rnx = (Rec<float>*)d.rec[nx];
    if (flags & F_NULL) {
        if (PHASE == 2) return;
        for (int64_t t = tr.t0; t < tr.t1; ++t) {
            const int64_t s = G.start + t * BS + lane;
            if (t * BS + lane < G.count) grid_insert(d, C, R, s, rin[s].x, rin[s].y, rin[s].z);
        }
        __syncthreads();
        prof_block(d, P_DRIFT, t_start, true, 0);
        return;
    }
    const bool split_a = (flags & F_SPLIT_A) != 0;
    const bool split_b = (flags & F_SPLIT_B) != 0;
    const bool drift = !FORCE_ONLY && !split_b;  // this launch moves the particles
    if (PHASE == 2 && split_a) return;            // no features before a split step's build
    if (PHASE == 1 && !drift) {
        // split step, second launch: its drifted records are in cur ^ 1
        for (int64_t t = tr.t0; t < tr.t1; ++t) {
            const int64_t s = G.start + t * BS + lane;
            if (t * BS + lane < G.count)
                d.mlp_pos[s] = make_float4((float)rnx[s].x, (float)rnx[s].y, (float)rnx[s].z, 0.f);
        }
        prof_block(d, P_DRIFT, t_start, true, 0);
        return;
    }
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double dt = P.dt, lam = C.lambda;
    const double m = G.mass;
    const double* gs = d.sums + (int64_t)d.blk_group[b] * NS;
    const double vr[3] = {sums_vref(gs, G.sum_m, 0), sums_vref(gs, G.sum_m, 1), sums_vref(gs, G.sum_m, 2)};
    const double c = P.cm[G.type];
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    bool oor = false;
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + lane;
        if (t * BS + lane >= G.count) {
            if (kFeat && !split_a) {
                const float z[kMlpK] = {};
                mlp_store_features(d, mc, s, z, false);  // padding slot
            }
            continue;
        }
        if (PHASE == 2) {
            float g[kMlpK];
            mlp_features_f32(d, mc, s, (float)bx, (float)by, (float)bz, g);
            oor |= mlp_store_features(d, mc, s, g, true);
            continue;
        }
        const Rec<float> a = rin[s];
        double x1 = a.x, y1 = a.y, z1 = a.z;
        if (kDrift && drift) {
            // integrator.rs:28-45
            const double vkx = a.vx + (double)a.fx * c;
            const double vky = a.vy + (double)a.fy * c;
            const double vkz = a.vz + (double)a.fz * c;
            part[S_KVX] += vkx * m;
            part[S_KVY] += vky * m;
            part[S_KVZ] += vkz * m;
            const double wx = vkx - vr[0], wy = vky - vr[1], wz = vkz - vr[2];
            part[S_KV2] += m * ((wx * wx + wy * wy) + wz * wz);
            const double v1x = vkx * lam, v1y = vky * lam, v1z = vkz * lam;
            x1 = wrap1(a.x + v1x * dt, bx);
            y1 = wrap1(a.y + v1y * dt, by);
            z1 = wrap1(a.z + v1z * dt, bz);
            part[S_VMAX] = fmax(part[S_VMAX], (v1x * v1x + v1y * v1y) + v1z * v1z);
            Rec<float> o = a;
            o.x = x1;
            o.y = y1;
            o.z = z1;
            o.vx = v1x;
            o.vy = v1y;
            o.vz = v1z;
            rnx[s] = o;
            d.mlp_pos[s] = make_float4((float)x1, (float)y1, (float)z1, 0.f);
            if (split_a) {
                grid_insert(d, C, R, s, x1, y1, z1);
                continue;
            }
        }
        if (!kFeat) continue;
        float g[kMlpK];
        mlp_features(d, mc, rin, s, x1, y1, z1, bx, by, bz, g);
        oor |= mlp_store_features(d, mc, s, g, true);
    }
    if (kFeat && oor) atomicOr(&d.mlp_stat[r].oor, 1u);
    if (kDrift && drift) partial_store<BS>(d, part, b, (P.th_kind == 2 ? kColsKick : 0u) | kColsDrift);
    if (PHASE == 2)
        prof_span_atomic(d, P_FEAT, t_start);
    else
        prof_block(d, P_DRIFT, t_start, true, 0);
}

// ---------------------------------------------------------------- network

// LDS row strides (BF16 elements), padded so that the 16 rows an MFMA
// fragment load touches fall on distinct banks.
constexpr int kStrK = kMlpK + 8;  // rows of 32
constexpr int kStrH = kMlpH + 8;  // rows of 128
constexpr int kNetWaves = 16;
constexpr int kNetBS = 64 * kNetWaves;

// D (16 x 16, FP32) of v_mfma_f32_16x16x32_bf16 in lane l: rows 4 (l / 16) + i,
// column l % 16.  A (16 x 32): lane l holds row l % 16, k = 8 (l / 16) .. + 7;
// B (32 x 16): lane l holds column l % 16, k = 8 (l / 16) .. + 7.
__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The network runs transposed: every layer computes (units x atoms) =
// W (units x k) . act^T (k x atoms), weights as the A operand from LDS and
// the activations as the B operand in registers.  A layer's D tile then
// leaves lane l holding atom l % 16 and units 16 t + 4 g + i (g = l / 16,
// i = 0..3), and the B operand of the next layer wants, for a K chunk c of
// 32 units, atom l % 16 and 8 units at K slots 8 g .. 8 g + 7.  Taking those
// 8 units from the lane's own values of tiles 2c and 2c + 1,
//   K slot 8 g + j  <->  unit 32 c + (j < 4 ? 4 g + j : 16 + 4 g + j - 4),
// keeps every activation in its lane (no LDS round trip between layers); the
// weights' K columns are staged in LDS in the same order, so each MFMA sums
// the same products.
__device__ __forceinline__ int net_perm(int k) {
    const int c = k & ~31, g = (k >> 3) & 3, j = k & 7;
    return c + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
}

// Forward and backward pass of 16 particles per wave and iteration.
__global__ __launch_bounds__(kNetBS) void k_mlp_net(DevState d) {
    __shared__ __attribute__((aligned(16))) __bf16 sW1[kMlpH * kStrK];   // W1 [n][k]: A of a1^T = W1 x^T
    __shared__ __attribute__((aligned(16))) __bf16 sW1t[kMlpK * kStrH];  // W1^T [k][perm n]: A of gx^T = W1^T d1^T
    __shared__ __attribute__((aligned(16))) __bf16 sW2[kMlpH * kStrH];   // W2 [m][perm n]: A of a2^T = W2 h1^T
    __shared__ __attribute__((aligned(16))) __bf16 sW2t[kMlpH * kStrH];  // W2^T [n][perm m]: A of t^T = W2^T d2^T
    __shared__ __attribute__((aligned(16))) float sb1[kMlpH];
    __shared__ __attribute__((aligned(16))) float sb2[kMlpH];
    __shared__ __attribute__((aligned(16))) float sw3[kMlpH];
    __shared__ __attribute__((aligned(16))) float sgs[kMlpK];  // e_scale * inv_std
    const uint64_t t_start = prof_now(d);
    const MlpConst& mc = *d.mlp_c;
    const __bf16* w = d.mlp_w;
    // rows of the BF16 weights into the padded LDS rows, 16 bytes at a time
    auto stage = [&](__bf16* dst, int stride, const __bf16* src, int rows, int cols) {
        const int per = cols / 8;
        for (int i = threadIdx.x; i < rows * per; i += kNetBS) {
            const int r = i / per, c8 = i % per;
            *(bf16x8*)(dst + r * stride + c8 * 8) = *(const bf16x8*)(src + r * cols + c8 * 8);
        }
    };
    // the same with the K columns (128 hidden units) in net_perm order, 8
    // bytes (4 units) at a time
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    auto stage_perm = [&](__bf16* dst, const __bf16* src, int rows) {
        constexpr int per = kMlpH / 4;
        for (int i = threadIdx.x; i < rows * per; i += kNetBS) {
            const int r = i / per, k4 = (i % per) * 4;
            *(bf16x4*)(dst + r * kStrH + k4) = *(const bf16x4*)(src + r * kMlpH + net_perm(k4));
        }
    };
    stage(sW1, kStrK, w + kMlpBfW1, kMlpH, kMlpK);
    stage_perm(sW1t, w + kMlpBfW1t, kMlpK);
    stage_perm(sW2, w + kMlpBfW2, kMlpH);
    stage_perm(sW2t, w + kMlpBfW2t, kMlpH);
    for (int i = threadIdx.x; i < kMlpH; i += kNetBS) {
        sb1[i] = d.mlp_theta[kMlpOffB1 + i];
        sb2[i] = d.mlp_theta[kMlpOffB2 + i];
        sw3[i] = d.mlp_theta[kMlpOffW3 + i];
    }
    for (int i = threadIdx.x; i < kMlpK; i += kNetBS) sgs[i] = mc.e_scale * mc.g_inv_std[i];
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int lr = l & 15, g = l >> 4;  // the lane's atom in the tile, its unit group
    const int64_t nrow = d.n_slots / 16;  // slots come in whole tiles of >= 64
    const int64_t rstep = (int64_t)gridDim.x * kNetWaves;
    // the features of the wave's next row tile load while this one runs
    int64_t rb = (int64_t)blockIdx.x * kNetWaves + wv;
    bf16x8 axn = {};
    if (rb < nrow) axn = *(const bf16x8*)(d.mlp_x + (rb * 16 + lr) * kMlpK + g * 8);
    for (; rb < nrow; rb += rstep) {
        // the weight fragments are read from LDS where they are used: kept
        // in registers across iterations they would need 320 VGPRs
        __asm__ volatile("" ::: "memory");
        const int64_t r0 = rb * 16;
        // a1^T = W1 x^T + b1 (K = 32 features: one MFMA per 16 units); the
        // lane's B operand is its atom's features 8 g .. 8 g + 7
        const bf16x8 ax = axn;
        if (rb + rstep < nrow) axn = *(const bf16x8*)(d.mlp_x + ((rb + rstep) * 16 + lr) * kMlpK + g * 8);
        float s1[8][4];  // silu'(a1) of units 16 t + 4 g + i
        bf16x8 h1[4];    // h1 as the B operand of K chunk c
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(*(const bf16x8*)(sW1 + (t * 16 + lr) * kStrK + g * 8), ax, acc);
            const f32x4 bb = *(const f32x4*)(sb1 + t * 16 + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float ds;
                const float h = silu_f(acc[i] + bb[i], ds);
                s1[t][i] = ds;
                h1[t / 2][(t & 1) * 4 + i] = (__bf16)h;
            }
        }
        // a2^T = W2 h1^T + b2, then y = w3 . h2 and d2 = w3 * silu'(a2)
        float y = 0.f;
        bf16x8 d2[4];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc = mfma(*(const bf16x8*)(sW2 + (t * 16 + lr) * kStrH + c * 32 + g * 8), h1[c], acc);
            const f32x4 bb = *(const f32x4*)(sb2 + t * 16 + 4 * g);
            const f32x4 ww = *(const f32x4*)(sw3 + t * 16 + 4 * g);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float ds;
                const float h = silu_f(acc[i] + bb[i], ds);
                y = fmaf(ww[i], h, y);
                d2[t / 2][(t & 1) * 4 + i] = (__bf16)(ww[i] * ds);
            }
        }
        // t^T = W2^T d2^T, d1 = t * silu'(a1)
        bf16x8 d1[4];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc = mfma(*(const bf16x8*)(sW2t + (t * 16 + lr) * kStrH + c * 32 + g * 8), d2[c], acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) d1[t / 2][(t & 1) * 4 + i] = (__bf16)(acc[i] * s1[t][i]);
        }
        // gx^T = W1^T d1^T; dE/dG = e_scale * inv_std * gx: the lane holds
        // features 16 t + 4 g .. + 3 of its atom
        const int64_t row = (r0 + lr) * kMlpK;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < 4; ++c)
                acc = mfma(*(const bf16x8*)(sW1t + (t * 16 + lr) * kStrH + c * 32 + g * 8), d1[c], acc);
            const f32x4 gs = *(const f32x4*)(sgs + t * 16 + 4 * g);
            f32x4 v;
            bf16x4 vb;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                v[i] = gs[i] * acc[i];
                vb[i] = (__bf16)v[i];
            }
            *(f32x4*)(d.mlp_gam + row + t * 16 + 4 * g) = v;
            *(bf16x4*)(d.mlp_gamb + row + t * 16 + 4 * g) = vb;
        }
        // per-atom energy: the four unit groups of each atom
        y += __shfl_xor(y, 16, 64);
        y += __shfl_xor(y, 32, 64);
        if (g == 0) d.mlp_e[r0 + lr] = mc.e_scale * (y + mc.b3) + mc.e_shift;
    }
    prof_span_atomic(d, P_NET, t_start);
}

// ---------------------------------------------------------------- forces

// Pair forces from dE/dG of both particles; LJ forces where the surrogate
// falls back and at the periodic check; then (stepping launches) the second
// half kick and barostat scale of integrator.rs:47-58 into buffer cur.
template <int BS, bool FORCE_ONLY>
__global__ __launch_bounds__(BS) void k_mlp_force(DevState d) {
    const int b = blockIdx.x;
    const GroupInfo G = d.groups[d.blk_group[b]];
    const int r = G.rep;
    const RepCtrl& C = ctrl_in(d, r);
    const uint32_t flags = FORCE_ONLY ? 0u : C.flags;
    if (flags & (F_EXIT | F_BUILD | F_NULL | F_SPLIT_A)) {
        prof_block(d, P_STEP, 0, false, 0);
        return;
    }
    const uint64_t t_start = prof_now(d);
    const MlpConst& mc = *d.mlp_c;
    const RepParams& P = d.params[r];
    const RepInfo R = d.reps[r];
    const TileRange tr = block_tiles(d, G, b, BS);
    const int lane = threadIdx.x;
    const int cur = C.cur;
    // stepping launches read the drifted state of cur ^ 1 and end in cur
    const Rec<float>* __restrict__ src = (const Rec<float>*)d.rec[FORCE_ONLY ? cur : (cur ^ 1)];
    Rec<float>* __restrict__ dst = (Rec<float>*)d.rec[cur];
    const bool fallback = d.mlp_stat[r].oor != 0u;
    const bool check = !FORCE_ONLY && mc.judge_every > 0 && (C.step + 1) % mc.judge_every == 0;
    const bool lj = fallback || check;
    const double bx = C.box[0], by = C.box[1], bz = C.box[2];
    const double mu = FORCE_ONLY ? 1.0 : C.myu;
    const double m = G.mass;
    const double* gs = d.sums + (int64_t)d.blk_group[b] * NS;
    const double vr[3] = {sums_vref(gs, G.sum_m, 0), sums_vref(gs, G.sum_m, 1), sums_vref(gs, G.sum_m, 2)};
    const double c = P.cm[G.type];
    const PairParam p0 = d.pairs[0];
    const double rc2 = (double)mc.rc * (double)mc.rc;
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    double jerr = 0.0, jref = 0.0;
    for (int64_t t = tr.t0; t < tr.t1; ++t) {
        const int64_t s = G.start + t * BS + lane;
        if (t * BS + lane >= G.count) continue;
        const Rec<float> a = src[s];
        float gi[kMlpK];
        {
            const f32x4* gp = (const f32x4*)(d.mlp_gam + s * kMlpK);
#pragma unroll
            for (int q = 0; q < kMlpK / 4; ++q) {
                const f32x4 v = gp[q];
                gi[4 * q] = v[0];
                gi[4 * q + 1] = v[1];
                gi[4 * q + 2] = v[2];
                gi[4 * q + 3] = v[3];
            }
        }
        float fx = 0.f, fy = 0.f, fz = 0.f, vir = 0.f;
        Acc<float> acc;  // LJ (fallback and check)
        const int nn = d.ncount[s];
        const int64_t ns = d.n_slots;
        // radial coefficient c of F_i += c d_ij / r_ij for a pair at rr < r_c:
        // c = sum_k G_k dg_k/dr with G = dE_i/dG + dE_j/dG and
        // dg_k/dr = e_k (-(r - mu_k) fc / w^2 + dfc/dr).  Two sums
        // A = sum G_k e_k and B = sum G_k e_k mu_k give
        // c = -(r A - B) fc / w^2 + A dfc/dr (three operations per k
        // besides the basis instead of five).
        auto coef = [&](float rr, const bf16x8 (&gj)[kMlpK / 8]) {
            const float arg = rr * mc.pi_rc;
            const float fc = 0.5f * (__cosf(arg) + 1.0f);
            const float dfc = -0.5f * mc.pi_rc * __sinf(arg);
            float sa = 0.f, sb = 0.f;
            gauss_basis(rr, mc, [&](int k, float e) {
                const float ge = (gi[k] + (float)gj[k / 8][k % 8]) * e;
                sa += ge;
                sb = fmaf(ge, mc.mu[k], sb);
            });
            return fmaf(-mc.inv_w2 * fc, fmaf(rr, sa, -sb), dfc * sa);
        };
        // software pipeline over the list: the index two entries ahead, the
        // position and dE/dG one entry ahead (both gathers go out together,
        // not after the distance test).  A neighbour's dE/dG is gathered in
        // BF16 (64 bytes instead of 128: the gathers bound this kernel); the
        // particle's own stays FP32.
        int32_t jn;  // index of the entry after the next
        bf16x8 gn[kMlpK / 8];
        auto gather_gam = [&](int32_t j) {
            const bf16x8* gp = (const bf16x8*)(d.mlp_gamb + (int64_t)j * kMlpK);
#pragma unroll
            for (int q8 = 0; q8 < kMlpK / 8; ++q8) gn[q8] = gp[q8];
        };
        if (FORCE_ONLY || lj) {
            // FP64 positions from the records: the LJ sums here are the LJ
            // path's own, bit for bit
            jn = nn > 1 ? d.nbr[ns + s] : 0;
            Rec<float> on;
            if (nn > 0) {
                const int32_t j0 = d.nbr[s];
                on = src[j0];
                gather_gam(j0);
            }
            for (int q = 0; q < nn; ++q) {
                const Rec<float> o = on;
                bf16x8 gj[kMlpK / 8];
#pragma unroll
                for (int q8 = 0; q8 < kMlpK / 8; ++q8) gj[q8] = gn[q8];
                if (q + 1 < nn) {
                    on = src[jn];
                    gather_gam(jn);
                    jn = q + 2 < nn ? d.nbr[(int64_t)(q + 2) * ns + s] : 0;
                }
                const double dx = min_image(o.x - a.x, bx);
                const double dy = min_image(o.y - a.y, by);
                const double dz = min_image(o.z - a.z, bz);
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                if (lj && r2 <= p0.r2max) acc.add(dx, dy, dz, r2, p0);
                if (!(r2 < rc2)) continue;
                const float rr = __builtin_amdgcn_sqrtf((float)r2);
                const float cc = coef(rr, gj);
                const float ir = __builtin_amdgcn_rcpf(rr);
                fx = fmaf(cc * ir, (float)dx, fx);
                fy = fmaf(cc * ir, (float)dy, fy);
                fz = fmaf(cc * ir, (float)dz, fz);
                vir = fmaf(cc, rr, vir);
            }
        } else {
            // surrogate-only steps: the entries within r_c that the features
            // kernel kept (in list order, so the sums are those of the whole
            // list), with the FP32 drifted positions it used (16 bytes per
            // neighbour)
            const float4* __restrict__ pos = d.mlp_pos;
            const int32_t* __restrict__ nb = d.mlp_nbr;
            const float4 pa = pos[s];
            const float bxf = (float)bx, byf = (float)by, bzf = (float)bz, rc2f = mc.rc * mc.rc;
            const int nc = d.mlp_ncount[s];
            jn = nc > 1 ? nb[ns + s] : 0;
            float4 on;
            if (nc > 0) {
                const int32_t j0 = nb[s];
                on = pos[j0];
                gather_gam(j0);
            }
            for (int q = 0; q < nc; ++q) {
                const float4 o = on;
                bf16x8 gj[kMlpK / 8];
#pragma unroll
                for (int q8 = 0; q8 < kMlpK / 8; ++q8) gj[q8] = gn[q8];
                if (q + 1 < nc) {
                    on = pos[jn];
                    gather_gam(jn);
                    jn = q + 2 < nc ? nb[(int64_t)(q + 2) * ns + s] : 0;
                }
                const float dx = min_image_f(o.x - pa.x, bxf);
                const float dy = min_image_f(o.y - pa.y, byf);
                const float dz = min_image_f(o.z - pa.z, bzf);
                const float r2 = (dx * dx + dy * dy) + dz * dz;
                if (!(r2 < rc2f)) continue;
                const float rr = __builtin_amdgcn_sqrtf(r2);
                const float cc = coef(rr, gj);
                const float ir = __builtin_amdgcn_rcpf(rr);
                fx = fmaf(cc * ir, dx, fx);
                fy = fmaf(cc * ir, dy, fy);
                fz = fmaf(cc * ir, dz, fz);
                vir = fmaf(cc, rr, vir);
            }
        }
        float pot = 2.0f * d.mlp_e[s];  // Particle::potential counts each energy twice
        if (check) {
            const double ex = (double)fx - (double)acc.fx, ey = (double)fy - (double)acc.fy,
                         ez = (double)fz - (double)acc.fz;
            jerr += (ex * ex + ey * ey) + ez * ez;
            jref += ((double)acc.fx * acc.fx + (double)acc.fy * acc.fy) + (double)acc.fz * acc.fz;
        }
        if (fallback) {
            fx = acc.fx;
            fy = acc.fy;
            fz = acc.fz;
            pot = acc.pot;
            vir = acc.vir;
        }
        Rec<float> out = a;
        if (!FORCE_ONLY) {
            out.vx = a.vx + (double)fx * c;
            out.vy = a.vy + (double)fy * c;
            out.vz = a.vz + (double)fz * c;
            out.x = a.x * mu;
            out.y = a.y * mu;
            out.z = a.z * mu;
        }
        out.fx = fx;
        out.fy = fy;
        out.fz = fz;
        dst[s] = out;
        if (FORCE_ONLY || (flags & F_WRITE_PV)) {
            d.pot[s] = (double)pot;
            d.vir[s] = (double)vir;
        }
        part[S_MVX] += out.vx * m;
        part[S_MVY] += out.vy * m;
        part[S_MVZ] += out.vz * m;
        {
            const double wx = out.vx - vr[0], wy = out.vy - vr[1], wz = out.vz - vr[2];
            part[S_MV2] += m * ((wx * wx + wy * wy) + wz * wz);
        }
        part[S_POT] += (double)pot;
        part[S_VIR] += (double)vir;
        part[S_WV] = fmax(part[S_WV], (out.vx * out.vx + out.vy * out.vy) + out.vz * out.vz);
        part[S_WA] = fmax(part[S_WA], (((double)fx * fx + (double)fy * fy) + (double)fz * fz) * P.inv_m2[G.type]);
        if (flags & F_INSERT) grid_insert(d, C, R, s, out.x, out.y, out.z);
    }
    partial_store<BS>(d, part, b, FORCE_ONLY ? kColsAll : kColsState);
    if (check) {
        double jv[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) jv[k] = 0.0;
        jv[0] = jerr;
        jv[1] = jref;
        block_store<BS>(jv, d.mlp_judge + b, d.n_blocks, 3u);
    }
    prof_block(d, P_STEP, t_start, true, 0);
}

// FP32 parameters -> BF16 inference copies (W1, W1^T, W2, W2^T).
__global__ __launch_bounds__(256) void k_mlp_convert(DevState d) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    __bf16* w = (__bf16*)d.mlp_w;
    const float* th = d.mlp_theta;
    if (i < kMlpH * kMlpK) {
        const int n = i / kMlpK, k = i % kMlpK;
        w[kMlpBfW1 + i] = (__bf16)th[kMlpOffW1 + i];
        w[kMlpBfW1t + k * kMlpH + n] = (__bf16)th[kMlpOffW1 + i];
    }
    if (i < kMlpH * kMlpH) {
        const int n = i / kMlpH, k = i % kMlpH;
        w[kMlpBfW2 + i] = (__bf16)th[kMlpOffW2 + i];
        w[kMlpBfW2t + k * kMlpH + n] = (__bf16)th[kMlpOffW2 + i];
    }
}

// ---------------------------------------------------------------- launchers

hipError_t launch_mlp_desc(const DevState& d, int bs, bool force_only, int phase, hipStream_t st) {
#define L(B)                                                                                              \
    if (force_only) hipLaunchKernelGGL((k_mlp_desc<B, true, 0>), dim3(d.n_blocks), dim3(B), 0, st, d);   \
    else if (phase == 1) hipLaunchKernelGGL((k_mlp_desc<B, false, 1>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else if (phase == 2) hipLaunchKernelGGL((k_mlp_desc<B, false, 2>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else return hipErrorInvalidValue;
    switch (bs) {
        case 64: L(64) break;
        case 128: L(128) break;
        case 256: L(256) break;
        case 512: L(512) break;
        default: return hipErrorInvalidValue;
    }
#undef L
    return hipGetLastError();
}

hipError_t launch_mlp_net(const DevState& d, int max_blocks, hipStream_t st) {
    // one block per CU (the weights fill most of its LDS): each block stages
    // the weights once and its waves stride over the row tiles
    const int64_t rows = d.n_slots / 16;
    int grid = (int)std::min<int64_t>((rows + kNetWaves - 1) / kNetWaves, std::max(1, max_blocks));
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_mlp_net, dim3(grid), dim3(kNetBS), 0, st, d);
    return hipGetLastError();
}

hipError_t launch_mlp_force(const DevState& d, int bs, bool force_only, hipStream_t st) {
#define L(B)                                                                                         \
    if (force_only) hipLaunchKernelGGL((k_mlp_force<B, true>), dim3(d.n_blocks), dim3(B), 0, st, d); \
    else hipLaunchKernelGGL((k_mlp_force<B, false>), dim3(d.n_blocks), dim3(B), 0, st, d);
    switch (bs) {
        case 64: L(64) break;
        case 128: L(128) break;
        case 256: L(256) break;
        case 512: L(512) break;
        default: return hipErrorInvalidValue;
    }
#undef L
    return hipGetLastError();
}

hipError_t launch_mlp_convert(const DevState& d, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_convert, dim3((kMlpH * kMlpH + 255) / 256), dim3(256), 0, st, d);
    return hipGetLastError();
}

}  // namespace md
