// md_mlp_train.hip -- training kernels of the learned per-atom energy
// surrogate (FP32), on frames of one periodic system size.
//
// Pair lists (all pairs within the cutoff, ascending j) and features are
// built once per data set.  A minibatch of M atoms (whole frames, taken in any
// order through a frame list) goes through
//   features   G,
//   forward    A1 = X W1^T, H1 = silu(A1 + b1), A2 = H1 W2^T, H2 = silu(A2 + b2),
//              y = H2 w3 + b3, E = e_scale y + e_shift,
//   backward   D2 = w3 * silu'(A2), T = D2 W2, D1 = T * silu'(A1), dE/dG = e_scale inv_std (D1 W1),
//   forces     F_i = sum_j (c_i + c_j)(r_ij) d_ij / r_ij (pair lists),
//   loss       L = w_e mean (E - E_ref)^2 + w_f mean (F - F_ref)^2,
//   gradient   the energy term by back-propagation and the force term by the
//              double backward through dE/dG:
//              U = d L / d(dE/dG) (pair kernel), V1 = U~ W1^T, PP = V1 silu'(A1),
//              V2 = PP W2^T, G2 = V2 w3 silu''(A2) + dy D2, G1 = (G2 W2) silu'(A1)
//              + T V1 silu''(A1), dW2 = G2^T H1 + D2^T PP, dW1 = G1^T X + D1^T U~,
//              db2 = sum G2, db1 = sum G1, dw3 = sum (V2 silu'(A2) + dy H2), db3 = sum dy,
//   update     Adam.
// Matrix products run on one tiled SGEMM kernel (NN, NT and TN forms, split-K
// partial sums added in a fixed order for the parameter gradients).
#include "md_common.h"
#include "md_train.h"

namespace md {

constexpr float kPiT = 3.14159265358979323846f;

__device__ __forceinline__ void silu3(float a, float& h, float& d1, float& d2) {
    const float s = 1.0f / (1.0f + expf(-a));
    h = a * s;
    d1 = s * (1.0f + a * (1.0f - s));
    d2 = s * (1.0f - s) * (2.0f + a * (1.0f - 2.0f * s));
}

// ---------------------------------------------------------------- SGEMM

// C[z] = alpha op(A) op(B) + beta C[z] over K range [z kc, (z + 1) kc);
// op(A) is M x K (TA: A stored K x M), op(B) is K x N (TB: B stored N x K).
constexpr int kGB = 64, kGK = 16;
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, int kc, float alpha, const float* __restrict__ A,
                                               int lda, const float* __restrict__ B, int ldb, float beta,
                                               float* __restrict__ C, int ldc, int64_t c_split) {
    __shared__ float As[kGK][kGB + 4];
    __shared__ float Bs[kGK][kGB + 4];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int m0 = blockIdx.y * kGB, n0 = blockIdx.x * kGB;
    const int k_lo = blockIdx.z * kc;
    const int k_hi = min(K, k_lo + kc);
    float acc[4][4] = {};
    for (int k0 = k_lo; k0 < k_hi; k0 += kGK) {
        for (int e = threadIdx.x; e < kGB * kGK; e += 256) {
            // A tile: element (m, k)
            int m, k;
            if (TA) {
                k = e / kGB;
                m = e % kGB;
            } else {
                m = e / kGK;
                k = e % kGK;
            }
            const int gm = m0 + m, gk = k0 + k;
            As[k][m] = (gm < M && gk < k_hi) ? (TA ? A[(int64_t)gk * lda + gm] : A[(int64_t)gm * lda + gk]) : 0.0f;
            int n, kb;
            if (TB) {
                n = e / kGK;
                kb = e % kGK;
            } else {
                kb = e / kGB;
                n = e % kGB;
            }
            const int gn = n0 + n, gkb = k0 + kb;
            Bs[kb][n] = (gn < N && gkb < k_hi) ? (TB ? B[(int64_t)gn * ldb + gkb] : B[(int64_t)gkb * ldb + gn]) : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kGK; ++k) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = As[k][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[k][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
    float* Cz = C + blockIdx.z * c_split;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int gm = m0 + ty + 16 * i;
        if (gm >= M) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int gn = n0 + tx + 16 * j;
            if (gn >= N) continue;
            float* cp = Cz + (int64_t)gm * ldc + gn;
            *cp = alpha * acc[i][j] + (beta != 0.0f ? beta * *cp : 0.0f);
        }
    }
}

// out[i] (+)= sum_z part[z][i] in split order
__global__ __launch_bounds__(256) void k_split_sum(const float* __restrict__ part, int splits, int64_t n,
                                                   float* __restrict__ out, int accumulate) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(int64_t)z * n + i];
    out[i] = accumulate ? out[i] + s : s;
}

// ---------------------------------------------------------------- pairs, features

// Positions of frame f staged in LDS 256 at a time; atom i visits every
// j != i of its frame in ascending order and calls visit(j) for the pairs with
// r^2 < rc^2 (the reference minimum image).
template <class Visit>
__device__ __forceinline__ void tr_all_pairs(const TrainState& t, int f, int i, Visit visit) {
    __shared__ double sx[256], sy[256], sz[256];
    const int n = t.n_atoms;
    const double* P = t.pos + (int64_t)f * n * 3;
    const double bx = t.box[3 * f], by = t.box[3 * f + 1], bz = t.box[3 * f + 2];
    double xi = 0, yi = 0, zi = 0;
    if (i < n) {
        xi = P[3 * i];
        yi = P[3 * i + 1];
        zi = P[3 * i + 2];
    }
    const double rc2 = (double)t.rc * t.rc;
    for (int j0 = 0; j0 < n; j0 += 256) {
        __syncthreads();
        const int jl = j0 + threadIdx.x;
        if (jl < n) {
            sx[threadIdx.x] = P[3 * jl];
            sy[threadIdx.x] = P[3 * jl + 1];
            sz[threadIdx.x] = P[3 * jl + 2];
        }
        __syncthreads();
        const int cnt = min(256, n - j0);
        if (i < n)
            for (int q = 0; q < cnt; ++q) {
                if (j0 + q == i) continue;
                const double dx = min_image(sx[q] - xi, bx), dy = min_image(sy[q] - yi, by),
                             dz = min_image(sz[q] - zi, bz);
                const double r2 = (dx * dx + dy * dy) + dz * dz;
                if (r2 < rc2) visit(j0 + q);
            }
    }
}

// neighbour counts of every atom of frames [f0, f0 + nf), and their maximum
__global__ __launch_bounds__(256) void k_tr_count(TrainState t, int f0, int nf, int32_t* max_cnt) {
    const int f = f0 + blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    int c = 0;
    tr_all_pairs(t, f, i, [&](int) { ++c; });
    if (i < t.n_atoms) {
        t.ncnt[(int64_t)f * t.n_atoms + i] = c;
        atomicMax(max_cnt, c);
    }
}

// neighbour indices (ascending) of every atom of frames [f0, f0 + nf)
__global__ __launch_bounds__(256) void k_tr_fill(TrainState t, int f0, int nf) {
    const int f = f0 + blockIdx.y;
    const int i = blockIdx.x * 256 + threadIdx.x;
    int32_t* out = t.nbr + ((int64_t)f * t.n_atoms + (i < t.n_atoms ? i : 0)) * t.cap;
    int c = 0;
    tr_all_pairs(t, f, i, [&](int j) {
        if (c < t.cap) out[c] = j;
        ++c;
    });
}

// displacement x_j - x_i (minimum image) and distance of a listed pair
struct TrPair {
    double dx, dy, dz;
    float r;
};
__device__ __forceinline__ TrPair tr_pair(const double* P, double xi, double yi, double zi, double bx, double by,
                                          double bz, int j) {
    TrPair p;
    p.dx = min_image(P[3 * j] - xi, bx);
    p.dy = min_image(P[3 * j + 1] - yi, by);
    p.dz = min_image(P[3 * j + 2] - zi, bz);
    p.r = (float)sqrt((p.dx * p.dx + p.dy * p.dy) + p.dz * p.dz);
    return p;
}

// G of every atom of frames [f0, f0 + nf) from the pair lists
__global__ __launch_bounds__(256) void k_tr_features(TrainState t, int f0, int nf) {
    const int f = f0 + blockIdx.y;
    const int n = t.n_atoms;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t row = (int64_t)f * n + i;
    const double* P = t.pos + (int64_t)f * n * 3;
    const double bx = t.box[3 * f], by = t.box[3 * f + 1], bz = t.box[3 * f + 2];
    const double xi = P[3 * i], yi = P[3 * i + 1], zi = P[3 * i + 2];
    float g[kMlpK];
#pragma unroll
    for (int k = 0; k < kMlpK; ++k) g[k] = 0.0f;
    const int cnt = t.ncnt[row];
    const int32_t* nb = t.nbr + row * t.cap;
    for (int q = 0; q < cnt; ++q) {
        const TrPair p = tr_pair(P, xi, yi, zi, bx, by, bz, nb[q]);
        const float fc = 0.5f * (cosf(kPiT * p.r / t.rc) + 1.0f);
#pragma unroll
        for (int k = 0; k < kMlpK; ++k) {
            const float u = p.r - t.mu[k];
            g[k] = fmaf(expf(-u * u * t.inv_2w2), fc, g[k]);
        }
    }
    float* out = t.G + row * kMlpK;
#pragma unroll
    for (int k = 0; k < kMlpK; ++k) out[k] = g[k];
}

// per-feature sum, sum of squares, min, max over all atoms (block partials)
__global__ __launch_bounds__(256) void k_tr_feature_stats(TrainState t, int64_t n_rows, double* part) {
    const int k = threadIdx.x % kMlpK, lane_row = threadIdx.x / kMlpK;  // 8 rows at a time
    double s = 0, s2 = 0, mn = 1e300, mx = -1e300;
    for (int64_t r = (int64_t)blockIdx.x * 8 + lane_row; r < n_rows; r += (int64_t)gridDim.x * 8) {
        const double v = t.G[r * kMlpK + k];
        s += v;
        s2 += v * v;
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    __shared__ double sh[4][256];
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = s2;
    sh[2][threadIdx.x] = mn;
    sh[3][threadIdx.x] = mx;
    __syncthreads();
    if (threadIdx.x < kMlpK) {
        for (int q = 1; q < 8; ++q) {
            const int o = q * kMlpK + threadIdx.x;
            s += sh[0][o];
            s2 += sh[1][o];
            mn = fmin(mn, sh[2][o]);
            mx = fmax(mx, sh[3][o]);
        }
        double* p = part + (int64_t)blockIdx.x * 4 * kMlpK;
        p[threadIdx.x] = s;
        p[kMlpK + threadIdx.x] = s2;
        p[2 * kMlpK + threadIdx.x] = mn;
        p[3 * kMlpK + threadIdx.x] = mx;
    }
}

// ---------------------------------------------------------------- elementwise

// X = (G - mean) inv_std for the M rows of the minibatch (slot b = frame bf[b])
__global__ __launch_bounds__(256) void k_tr_normalize(TrainState t, int64_t M) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= M * kMlpK) return;
    const int k = (int)(e % kMlpK);
    const int64_t r = e / kMlpK;
    const int64_t b = r / t.n_atoms, i = r % t.n_atoms;
    const int64_t src = ((int64_t)t.bf[b] * t.n_atoms + i) * kMlpK + k;
    t.X[e] = (t.G[src] - t.g_mean[k]) * t.g_inv_std[k];
}

// layer activations: H = silu(A + b), S = silu', Sp = silu'' (in place over A -> H)
__global__ __launch_bounds__(256) void k_tr_act(int64_t M, const float* __restrict__ b, float* __restrict__ AH,
                                                float* __restrict__ S, float* __restrict__ Sp) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= M * kMlpH) return;
    float h, d1, d2;
    silu3(AH[e] + b[e % kMlpH], h, d1, d2);
    AH[e] = h;
    S[e] = d1;
    Sp[e] = d2;
}

// per row: y = H2 . w3 + b3, E = e_scale y + e_shift; D2 = w3 * S2
__global__ __launch_bounds__(256) void k_tr_head(TrainState t, int64_t M) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= M) return;
    const float* w3 = t.theta + kMlpOffW3;
    float y = t.theta[kMlpOffB3];
    for (int m = 0; m < kMlpH; ++m) {
        y = fmaf(t.H2[r * kMlpH + m], w3[m], y);
        t.D2[r * kMlpH + m] = w3[m] * t.S2[r * kMlpH + m];
    }
    t.E[r] = t.e_scale * y + t.e_shift;
}

// D1 = T * S1
__global__ __launch_bounds__(256) void k_tr_mul(int64_t n, const float* __restrict__ a, const float* __restrict__ b,
                                                float* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) out[e] = a[e] * b[e];
}

// Gam = e_scale inv_std GX (in place), Ut likewise for U
__global__ __launch_bounds__(256) void k_tr_scale_in(TrainState t, int64_t M, float* __restrict__ v) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= M * kMlpK) return;
    v[e] *= t.e_scale * t.g_inv_std[e % kMlpK];
}

// G2 = V2 w3 S2p + dy D2 (V2 in place -> G2), and the w3 gradient terms
// W3T = V2 S2 + dy H2 (row-wise, summed later)
__global__ __launch_bounds__(256) void k_tr_g2(TrainState t, int64_t M) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= M * kMlpH) return;
    const int m = (int)(e % kMlpH);
    const int64_t r = e / kMlpH;
    const float v2 = t.V2[e], dy = t.dy[r];
    t.W3T[e] = v2 * t.S2[e] + dy * t.H2[e];
    t.V2[e] = v2 * t.theta[kMlpOffW3 + m] * t.S2p[e] + dy * t.D2[e];
}

// G1 = R1a * S1 + T * V1 * S1p (R1a in place -> G1); PP = V1 * S1 computed earlier
__global__ __launch_bounds__(256) void k_tr_g1(TrainState t, int64_t M) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= M * kMlpH) return;
    t.R1[e] = t.R1[e] * t.S1[e] + t.T[e] * t.V1[e] * t.S1p[e];
}

// ---------------------------------------------------------------- pair kernels

// Forces of the minibatch frames from Gam (rows r = b n + i), the loss terms
// and their derivatives phi = dL/dF, dy = dL/dy.
__global__ __launch_bounds__(256) void k_tr_forces(TrainState t, double* loss_part) {
    const int b = blockIdx.y, f = t.bf[b];
    const int n = t.n_atoms;
    const int i = blockIdx.x * 256 + threadIdx.x;
    double le = 0.0, lf = 0.0;
    if (i < n) {
        const int64_t ri = (int64_t)b * n + i;  // row in the minibatch
        const int64_t di = (int64_t)f * n + i;  // row in the data set
        const double* P = t.pos + (int64_t)f * n * 3;
        const double bx = t.box[3 * f], by = t.box[3 * f + 1], bz = t.box[3 * f + 2];
        const double xi = P[3 * i], yi = P[3 * i + 1], zi = P[3 * i + 2];
        float gi[kMlpK];
#pragma unroll
        for (int k = 0; k < kMlpK; ++k) gi[k] = t.Gam[ri * kMlpK + k];
        float fx = 0.f, fy = 0.f, fz = 0.f;
        const int cnt = t.ncnt[di];
        const int32_t* nb = t.nbr + di * t.cap;
        for (int q = 0; q < cnt; ++q) {
            const int j = nb[q];
            const TrPair p = tr_pair(P, xi, yi, zi, bx, by, bz, j);
            const float a = kPiT * p.r / t.rc;
            const float fc = 0.5f * (cosf(a) + 1.0f), dfc = -0.5f * kPiT / t.rc * sinf(a);
            const float* gj = t.Gam + ((int64_t)b * n + j) * kMlpK;
            float c = 0.f;
#pragma unroll
            for (int k = 0; k < kMlpK; ++k) {
                const float u = p.r - t.mu[k];
                const float e = expf(-u * u * t.inv_2w2);
                c = fmaf(gi[k] + gj[k], e * fmaf(-u * t.inv_w2, fc, dfc), c);
            }
            const float s = c / p.r;
            fx = fmaf(s, (float)p.dx, fx);
            fy = fmaf(s, (float)p.dy, fy);
            fz = fmaf(s, (float)p.dz, fz);
        }
        const float ex = fx - t.Fref[3 * di], ey = fy - t.Fref[3 * di + 1], ez = fz - t.Fref[3 * di + 2];
        const float ee = t.E[ri] - t.Eref[di];
        t.F[3 * ri] = fx;
        t.F[3 * ri + 1] = fy;
        t.F[3 * ri + 2] = fz;
        const float cf = 2.0f * t.w_f / (3.0f * (float)t.n_total);
        t.phi[3 * ri] = cf * ex;
        t.phi[3 * ri + 1] = cf * ey;
        t.phi[3 * ri + 2] = cf * ez;
        t.dy[ri] = 2.0f * t.w_e / (float)t.n_total * ee * t.e_scale;
        le = (double)ee * ee;
        lf = ((double)ex * ex + (double)ey * ey) + (double)ez * ez;
    }
    // block sums of the squared errors (fixed order)
    __shared__ double sl[2][256];
    sl[0][threadIdx.x] = le;
    sl[1][threadIdx.x] = lf;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0, sb = 0;
        for (int q = 0; q < 256; ++q) {
            sa += sl[0][q];
            sb += sl[1][q];
        }
        const int64_t blk = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
        loss_part[2 * blk] = sa;
        loss_part[2 * blk + 1] = sb;
    }
}

// out[0..1] = block partials of the squared errors summed in block order
__global__ __launch_bounds__(64) void k_tr_loss_sum(const double* __restrict__ part, int64_t blocks,
                                                   double* __restrict__ out) {
    if (threadIdx.x >= 2) return;
    double s = 0.0;
    for (int64_t q = 0; q < blocks; ++q) s += part[2 * q + threadIdx.x];
    out[threadIdx.x] = s;
}

// U_i = sum_j ((phi_i - phi_j) . d_ij / r_ij) dg(r_ij)
__global__ __launch_bounds__(256) void k_tr_u(TrainState t) {
    const int b = blockIdx.y, f = t.bf[b];
    const int n = t.n_atoms;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t ri = (int64_t)b * n + i;
    const int64_t di = (int64_t)f * n + i;
    const double* P = t.pos + (int64_t)f * n * 3;
    const double bx = t.box[3 * f], by = t.box[3 * f + 1], bz = t.box[3 * f + 2];
    const double xi = P[3 * i], yi = P[3 * i + 1], zi = P[3 * i + 2];
    const float px = t.phi[3 * ri], py = t.phi[3 * ri + 1], pz = t.phi[3 * ri + 2];
    float u[kMlpK];
#pragma unroll
    for (int k = 0; k < kMlpK; ++k) u[k] = 0.0f;
    const int cnt = t.ncnt[di];
    const int32_t* nb = t.nbr + di * t.cap;
    for (int q = 0; q < cnt; ++q) {
        const int j = nb[q];
        const TrPair p = tr_pair(P, xi, yi, zi, bx, by, bz, j);
        const float* pj = t.phi + 3 * ((int64_t)b * n + j);
        const float proj = ((px - pj[0]) * (float)p.dx + (py - pj[1]) * (float)p.dy + (pz - pj[2]) * (float)p.dz) / p.r;
        const float a = kPiT * p.r / t.rc;
        const float fc = 0.5f * (cosf(a) + 1.0f), dfc = -0.5f * kPiT / t.rc * sinf(a);
#pragma unroll
        for (int k = 0; k < kMlpK; ++k) {
            const float w = p.r - t.mu[k];
            const float e = expf(-w * w * t.inv_2w2);
            u[k] = fmaf(proj, e * fmaf(-w * t.inv_w2, fc, dfc), u[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < kMlpK; ++k) t.U[ri * kMlpK + k] = u[k];
}

// column sums of an M x ncol matrix into out[0..ncol) (block partials, fixed order)
__global__ __launch_bounds__(256) void k_tr_colsum(const float* __restrict__ A, int64_t M, int ncol,
                                                   float* __restrict__ part) {
    // block b sums rows [b * rows_per, ...) for column threadIdx.x (ncol <= 256)
    const int c = threadIdx.x;
    if (c >= ncol) return;
    const int64_t per = (M + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = min(M, r0 + per);
    float s = 0.0f;
    for (int64_t r = r0; r < r1; ++r) s += A[r * ncol + c];
    part[(int64_t)blockIdx.x * ncol + c] = s;
}

// Adam on the flat parameter vector; the step count t is 1-based.
__global__ __launch_bounds__(256) void k_tr_adam(float* __restrict__ theta, const float* __restrict__ g,
                                                 float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                                 float b1, float b2, float eps, int step) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)step));
    const float vh = vi / (1.0f - powf(b2, (float)step));
    theta[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------- launchers

static inline unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t tr_gemm(bool ta, bool tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B,
                   int ldb, float beta, float* C, int ldc, int splits, hipStream_t st) {
    if (splits < 1) splits = 1;
    const int kc = ((K + splits - 1) / splits + kGK - 1) / kGK * kGK;
    splits = (K + kc - 1) / kc;
    const dim3 grid((N + kGB - 1) / kGB, (M + kGB - 1) / kGB, splits);
    const int64_t cs = (int64_t)M * ldc;
    if (!ta && !tb) hipLaunchKernelGGL((k_sgemm<false, false>), grid, dim3(256), 0, st, M, N, K, kc, alpha, A, lda, B, ldb, beta, C, ldc, cs);
    if (!ta && tb) hipLaunchKernelGGL((k_sgemm<false, true>), grid, dim3(256), 0, st, M, N, K, kc, alpha, A, lda, B, ldb, beta, C, ldc, cs);
    if (ta && !tb) hipLaunchKernelGGL((k_sgemm<true, false>), grid, dim3(256), 0, st, M, N, K, kc, alpha, A, lda, B, ldb, beta, C, ldc, cs);
    if (ta && tb) hipLaunchKernelGGL((k_sgemm<true, true>), grid, dim3(256), 0, st, M, N, K, kc, alpha, A, lda, B, ldb, beta, C, ldc, cs);
    return hipGetLastError();
}

int tr_gemm_splits(int K, int splits) {
    const int kc = ((K + splits - 1) / splits + kGK - 1) / kGK * kGK;
    return (K + kc - 1) / kc;
}

hipError_t tr_split_sum(const float* part, int splits, int64_t n, float* out, bool accumulate, hipStream_t st) {
    hipLaunchKernelGGL(k_split_sum, dim3(nblk(n)), dim3(256), 0, st, part, splits, n, out, accumulate ? 1 : 0);
    return hipGetLastError();
}

hipError_t tr_count_pairs(const TrainState& t, int f0, int nf, int32_t* max_cnt, hipStream_t st) {
    hipLaunchKernelGGL(k_tr_count, dim3(nblk(t.n_atoms), nf), dim3(256), 0, st, t, f0, nf, max_cnt);
    return hipGetLastError();
}

hipError_t tr_fill_pairs(const TrainState& t, int f0, int nf, hipStream_t st) {
    hipLaunchKernelGGL(k_tr_fill, dim3(nblk(t.n_atoms), nf), dim3(256), 0, st, t, f0, nf);
    return hipGetLastError();
}

hipError_t tr_features(const TrainState& t, int f0, int nf, hipStream_t st) {
    hipLaunchKernelGGL(k_tr_features, dim3(nblk(t.n_atoms), nf), dim3(256), 0, st, t, f0, nf);
    return hipGetLastError();
}

hipError_t tr_feature_stats(const TrainState& t, int64_t rows, double* part, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(k_tr_feature_stats, dim3(blocks), dim3(256), 0, st, t, rows, part);
    return hipGetLastError();
}

hipError_t tr_forward_backward(const TrainState& t, int nf, hipStream_t st) {
    hipError_t e;
    const float* th = t.theta;
    const int64_t M = (int64_t)nf * t.n_atoms;
    hipLaunchKernelGGL(k_tr_normalize, dim3(nblk(M * kMlpK)), dim3(256), 0, st, t, M);
    // A1 = X W1^T -> H1
    if ((e = tr_gemm(false, true, (int)M, kMlpH, kMlpK, 1.f, t.X, kMlpK, th + kMlpOffW1, kMlpK, 0.f, t.H1, kMlpH, 1, st))) return e;
    hipLaunchKernelGGL(k_tr_act, dim3(nblk(M * kMlpH)), dim3(256), 0, st, M, th + kMlpOffB1, t.H1, t.S1, t.S1p);
    // A2 = H1 W2^T -> H2
    if ((e = tr_gemm(false, true, (int)M, kMlpH, kMlpH, 1.f, t.H1, kMlpH, th + kMlpOffW2, kMlpH, 0.f, t.H2, kMlpH, 1, st))) return e;
    hipLaunchKernelGGL(k_tr_act, dim3(nblk(M * kMlpH)), dim3(256), 0, st, M, th + kMlpOffB2, t.H2, t.S2, t.S2p);
    hipLaunchKernelGGL(k_tr_head, dim3(nblk(M)), dim3(256), 0, st, t, M);
    // T = D2 W2, D1 = T S1, Gam = e_scale inv_std (D1 W1)
    if ((e = tr_gemm(false, false, (int)M, kMlpH, kMlpH, 1.f, t.D2, kMlpH, th + kMlpOffW2, kMlpH, 0.f, t.T, kMlpH, 1, st))) return e;
    hipLaunchKernelGGL(k_tr_mul, dim3(nblk(M * kMlpH)), dim3(256), 0, st, M * kMlpH, t.T, t.S1, t.D1);
    if ((e = tr_gemm(false, false, (int)M, kMlpK, kMlpH, 1.f, t.D1, kMlpH, th + kMlpOffW1, kMlpK, 0.f, t.Gam, kMlpK, 1, st))) return e;
    hipLaunchKernelGGL(k_tr_scale_in, dim3(nblk(M * kMlpK)), dim3(256), 0, st, t, M, t.Gam);
    return hipGetLastError();
}

hipError_t tr_forces(const TrainState& t, int nf, double* loss_part, hipStream_t st) {
    hipLaunchKernelGGL(k_tr_forces, dim3(nblk(t.n_atoms), nf), dim3(256), 0, st, t, loss_part);
    return hipGetLastError();
}

hipError_t tr_loss_sum(const TrainState& t, int nf, const double* loss_part, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_tr_loss_sum, dim3(1), dim3(64), 0, st, loss_part, (int64_t)nblk(t.n_atoms) * nf, out);
    return hipGetLastError();
}

// Gradient of the minibatch loss added into t.grad (which the caller clears).
hipError_t tr_gradient(const TrainState& t, int nf, hipStream_t st) {
    hipError_t e;
    const float* th = t.theta;
    const int64_t M = (int64_t)nf * t.n_atoms;
    hipLaunchKernelGGL(k_tr_u, dim3(nblk(t.n_atoms), nf), dim3(256), 0, st, t);
    hipLaunchKernelGGL(k_tr_scale_in, dim3(nblk(M * kMlpK)), dim3(256), 0, st, t, M, t.U);
    // V1 = Ut W1^T, PP = V1 S1, V2 = PP W2^T
    if ((e = tr_gemm(false, true, (int)M, kMlpH, kMlpK, 1.f, t.U, kMlpK, th + kMlpOffW1, kMlpK, 0.f, t.V1, kMlpH, 1, st))) return e;
    hipLaunchKernelGGL(k_tr_mul, dim3(nblk(M * kMlpH)), dim3(256), 0, st, M * kMlpH, t.V1, t.S1, t.PP);
    if ((e = tr_gemm(false, true, (int)M, kMlpH, kMlpH, 1.f, t.PP, kMlpH, th + kMlpOffW2, kMlpH, 0.f, t.V2, kMlpH, 1, st))) return e;
    // G2 (in V2) and the w3 terms
    hipLaunchKernelGGL(k_tr_g2, dim3(nblk(M * kMlpH)), dim3(256), 0, st, t, M);
    // R1 = G2 W2, G1 (in R1)
    if ((e = tr_gemm(false, false, (int)M, kMlpH, kMlpH, 1.f, t.V2, kMlpH, th + kMlpOffW2, kMlpH, 0.f, t.R1, kMlpH, 1, st))) return e;
    hipLaunchKernelGGL(k_tr_g1, dim3(nblk(M * kMlpH)), dim3(256), 0, st, t, M);
    // parameter gradients: split-K over the atoms, partial sums in split order
    const int S = t.splits;
    float* g = t.grad;
    const int sW2 = tr_gemm_splits((int)M, S);
    if ((e = tr_gemm(true, false, kMlpH, kMlpH, (int)M, 1.f, t.V2, kMlpH, t.H1, kMlpH, 0.f, t.gpart, kMlpH, S, st))) return e;
    if ((e = tr_split_sum(t.gpart, sW2, kMlpH * kMlpH, g + kMlpOffW2, true, st))) return e;
    if ((e = tr_gemm(true, false, kMlpH, kMlpH, (int)M, 1.f, t.D2, kMlpH, t.PP, kMlpH, 0.f, t.gpart, kMlpH, S, st))) return e;
    if ((e = tr_split_sum(t.gpart, sW2, kMlpH * kMlpH, g + kMlpOffW2, true, st))) return e;
    if ((e = tr_gemm(true, false, kMlpH, kMlpK, (int)M, 1.f, t.R1, kMlpH, t.X, kMlpK, 0.f, t.gpart, kMlpK, S, st))) return e;
    if ((e = tr_split_sum(t.gpart, sW2, kMlpH * kMlpK, g + kMlpOffW1, true, st))) return e;
    if ((e = tr_gemm(true, false, kMlpH, kMlpK, (int)M, 1.f, t.D1, kMlpH, t.U, kMlpK, 0.f, t.gpart, kMlpK, S, st))) return e;
    if ((e = tr_split_sum(t.gpart, sW2, kMlpH * kMlpK, g + kMlpOffW1, true, st))) return e;
    // bias and w3 gradients: column sums
    const int cb = t.colsum_blocks;
    hipLaunchKernelGGL(k_tr_colsum, dim3(cb), dim3(256), 0, st, t.V2, M, kMlpH, t.gpart);
    if ((e = tr_split_sum(t.gpart, cb, kMlpH, g + kMlpOffB2, true, st))) return e;
    hipLaunchKernelGGL(k_tr_colsum, dim3(cb), dim3(256), 0, st, t.R1, M, kMlpH, t.gpart);
    if ((e = tr_split_sum(t.gpart, cb, kMlpH, g + kMlpOffB1, true, st))) return e;
    hipLaunchKernelGGL(k_tr_colsum, dim3(cb), dim3(256), 0, st, t.W3T, M, kMlpH, t.gpart);
    if ((e = tr_split_sum(t.gpart, cb, kMlpH, g + kMlpOffW3, true, st))) return e;
    hipLaunchKernelGGL(k_tr_colsum, dim3(cb), dim3(256), 0, st, t.dy, M, 1, t.gpart);
    if ((e = tr_split_sum(t.gpart, cb, 1, g + kMlpOffB3, true, st))) return e;
    return hipGetLastError();
}

hipError_t tr_adam(const TrainState& t, float lr, int step, hipStream_t st) {
    hipLaunchKernelGGL(k_tr_adam, dim3(nblk(kMlpNP)), dim3(256), 0, st, t.theta, t.grad, t.adam_m, t.adam_v, kMlpNP, lr,
                       0.9f, 0.999f, 1e-8f, step);
    return hipGetLastError();
}

}  // namespace md
