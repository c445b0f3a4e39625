// md_common.h -- device helpers shared by the kernel files of the MD step
// engine: the reference's minimum image and wrap, pair terms, wave / block
// reductions, cell-grid insertion and the per-launch profiling records.
#pragma once

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

// FP32 minimum image of one component (the list build's prefilter, the
// surrogate's FP32 distances).
__device__ __forceinline__ float min_image_f(float r, float b) {
    // selects, not branches: the candidate loop stays one basic block
    const float h = 0.5f * b;
    return r < -h ? r + b : (r > h ? r - b : r);
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

// The GPU deals blocks to its 8 XCDs in turn (block b to XCD b % 8), and each
// XCD has its own L2.  xcd_block gives physical block b the logical block it
// works on such that each XCD runs one contiguous range of logical blocks:
// blocks of neighbouring particles then share an L2.  A bijection on [0, n).
__device__ __forceinline__ int xcd_block(int b, int n) {
    constexpr int X = 8;
    const int q = n / X, r = n % X;
    const int x = b % X;
    return x * q + (x < r ? x : r) + b / X;
}

// A value the whole wave holds alike, moved to scalar registers (it was
// loaded through the vector path, the compiler cannot tell it is uniform).
__device__ __forceinline__ double wave_uniform(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ float wave_uniform(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ PairParam wave_uniform(const PairParam& p) {
    PairParam q;
    q.sigma = wave_uniform(p.sigma);
    q.eps = wave_uniform(p.eps);
    q.r_cut = wave_uniform(p.r_cut);
    q.u_cut = wave_uniform(p.u_cut);
    q.r2max = wave_uniform(p.r2max);
    q.sig2f = wave_uniform(p.sig2f);
    q.eps24f = wave_uniform(p.eps24f);
    q.eps4f = wave_uniform(p.eps4f);
    q.ucutf = wave_uniform(p.ucutf);
    return q;
}

// Barrier for LDS traffic only: every thread's LDS accesses before it are
// complete and visible to the block after it.  Unlike __syncthreads it does
// not wait for outstanding global loads and stores, which stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// DPP move of a double from the lane that CTRL selects; lanes without a
// source lane, or in rows that ROW_MASK leaves out, get `fill`.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v, double fill) {
    const uint64_t b = __builtin_bit_cast(uint64_t, v), f = __builtin_bit_cast(uint64_t, fill);
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)b, CTRL, ROW_MASK, 0xf, false);
    const int hi =
        __builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32), (int)(uint32_t)(b >> 32), CTRL, ROW_MASK, 0xf, false);
    return __builtin_bit_cast(double, (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32));
}

// Inclusive scan of a wave's 64 values in a fixed tree on the VALU: shifts by
// 1, 2, 4 and 8 lanes within each row of 16, then row 0's last lane added to
// row 1 and row 2's to row 3, then lane 31 added to rows 2 and 3.  Lane 63
// ends with the sum (MAX: the maximum) of all 64 values.
template <bool MAX>
__device__ __forceinline__ double wave_total(double v) {
    const double id = MAX ? -__builtin_inf() : 0.0;
    auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
    v = op(v, dpp_f64<0x111, 0xf>(v, id));  // row_shr:1
    v = op(v, dpp_f64<0x112, 0xf>(v, id));  // row_shr:2
    v = op(v, dpp_f64<0x114, 0xf>(v, id));  // row_shr:4
    v = op(v, dpp_f64<0x118, 0xf>(v, id));  // row_shr:8
    v = op(v, dpp_f64<0x142, 0xa>(v, id));  // row_bcast:15 into rows 1, 3
    v = op(v, dpp_f64<0x143, 0xc>(v, id));  // row_bcast:31 into rows 2, 3
    return v;
}

// Block reduction of NS columns in a fixed order (wave scan tree, then waves
// in index order); thread k < NS receives column k's total when mask has it
// (returns false otherwise).  It returns after the closing barrier, so no
// thread waits for the stores a caller makes of the totals.
template <int BS>
__device__ __forceinline__ bool block_reduce(const double (&v)[NS], uint32_t mask, double& total) {
    constexpr int NW = BS / 64;
    __shared__ double red[NW * NS];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        if (!((mask >> k) & 1u)) continue;
        const double t = k >= kFirstMax ? wave_total<true>(v[k]) : wave_total<false>(v[k]);
        if (lane == 63) red[w * NS + k] = t;
    }
    lds_barrier();
    const int k = threadIdx.x;
    const bool mine = k < NS && ((mask >> k) & 1u);
    double acc = 0.0;
    if (mine) {
        acc = red[k];
        for (int i = 1; i < NW; ++i)
            acc = k >= kFirstMax ? fmax(acc, red[i * NS + k]) : acc + red[i * NS + k];
    }
    lds_barrier();  // red may be reused by the next call
    total = acc;
    return mine;
}

// block_reduce, column k stored to out[k * stride].
template <int BS>
__device__ __forceinline__ void block_store(const double (&v)[NS], double* out, int64_t stride,
                                            uint32_t mask) {
    double t;
    if (block_reduce<BS>(v, mask, t)) out[threadIdx.x * stride] = t;
}

// ---------------------------------------------------------------- exact accumulators

// v in units of 2^-60, rounded to nearest (ties to even), as three signed
// 40-bit digits d0 + d1 2^40 + d2 2^80.  Returns false when v is not finite
// or |v| >= 2^60.
__device__ __forceinline__ bool acc_digits(double v, long long (&dg)[kAccDigits]) {
    dg[0] = dg[1] = dg[2] = 0;
    if (v == 0.0) return true;
    const double a = fabs(v);
    if (!(a < 0x1p60)) return false;  // also NaN
    int e;
    const double m = frexp(a, &e);                          // a = m 2^e, m in [0.5, 1)
    unsigned long long M = (unsigned long long)(m * 0x1p53);  // exact: a = M 2^(e - 53)
    int sh = e - 53 + 60;                                    // a = M 2^sh units
    if (sh < 0) {
        const int r = -sh;
        if (r > 54) return true;  // below half a unit
        const unsigned long long q = M >> r, rem = M - (q << r), half = 1ull << (r - 1);
        M = q + ((rem > half || (rem == half && (q & 1ull))) ? 1ull : 0ull);
        sh = 0;
    }
    // N = M << sh, sh <= 67: 128 bits as hi:lo
    unsigned long long lo, hi;
    if (sh == 0) {
        lo = M;
        hi = 0;
    } else if (sh < 64) {
        lo = M << sh;
        hi = M >> (64 - sh);
    } else {
        lo = 0;
        hi = M << (sh - 64);
    }
    constexpr unsigned long long mask = (1ull << 40) - 1ull;
    const long long s = v < 0.0 ? -1 : 1;
    dg[0] = s * (long long)(lo & mask);
    dg[1] = s * (long long)(((lo >> 40) | (hi << 24)) & mask);
    dg[2] = s * (long long)((hi >> 16) & mask);
    return true;
}

// Column k's partial of logical block b into the accumulators.
__device__ __forceinline__ void acc_add(const DevState& d, int b, int k, double v) {
    const int seg = (int)(((int64_t)b * kAccSeg) / d.n_blocks);
    unsigned long long* w = d.acc + ((int64_t)seg * NS + k) * kAccDigits;
    if (k >= kFirstMax) {
        // non-negative maxima order like their bit patterns; NaN is skipped as fmax skips it
        if (!isnan(v)) __hip_atomic_fetch_max(w, __builtin_bit_cast(unsigned long long, v), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    long long dg[kAccDigits];
    if (!acc_digits(v, dg)) {
        __hip_atomic_fetch_or(d.acc + (int64_t)kAccSeg * NS * kAccDigits + k, 1ull, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
#pragma unroll
    for (int q = 0; q < kAccDigits; ++q)
        if (dg[q] != 0)
            __hip_atomic_fetch_add(w + q, (unsigned long long)dg[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Column k's total from the accumulators (the mark makes it NaN).
__device__ __forceinline__ double acc_total(const DevState& d, int k) {
    const unsigned long long* base = d.acc;
    if (k >= kFirstMax) {
        unsigned long long mx = 0;
#pragma unroll
        for (int seg = 0; seg < kAccSeg; ++seg) {
            const unsigned long long x = base[((int64_t)seg * NS + k) * kAccDigits];
            mx = x > mx ? x : mx;
        }
        return __builtin_bit_cast(double, mx);
    }
    long long s[kAccDigits] = {0, 0, 0};
#pragma unroll
    for (int seg = 0; seg < kAccSeg; ++seg) {
        const unsigned long long* w = base + ((int64_t)seg * NS + k) * kAccDigits;
#pragma unroll
        for (int q = 0; q < kAccDigits; ++q) s[q] += (long long)w[q];
    }
    if (base[(int64_t)kAccSeg * NS * kAccDigits + k] != 0ull) return __builtin_nan("");
    return ((double)s[2] * 0x1p20 + (double)s[1] * 0x1p-20) + (double)s[0] * 0x1p-60;
}

// Column k's accumulator words reset for the next launch.
__device__ __forceinline__ void acc_clear(const DevState& d, int k) {
#pragma unroll
    for (int seg = 0; seg < kAccSeg; ++seg)
#pragma unroll
        for (int q = 0; q < kAccDigits; ++q) d.acc[((int64_t)seg * NS + k) * kAccDigits + q] = 0ull;
    d.acc[(int64_t)kAccSeg * NS * kAccDigits + k] = 0ull;
}

// The partials of logical block b: a row of `partial`, or the accumulators.
template <int BS>
__device__ __forceinline__ void partial_store(const DevState& d, const double (&v)[NS], int b, uint32_t mask) {
    double t;
    if (!block_reduce<BS>(v, mask, t)) return;
    if (d.acc_mode)
        acc_add(d, b, threadIdx.x, t);
    else
        d.partial[(int64_t)threadIdx.x * d.n_blocks + b] = t;
}

// In-place exclusive scan of a[0..n) (n <= kBrickCells); a[n] receives the total.
template <int BS>
__device__ __forceinline__ void block_exclusive_scan(int* a, int n) {
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
    lds_barrier();
    int before = 0;
    for (int i = 0; i < w; ++i) before += wsum[i];
    int run = before + incl - local;
    for (int i = lo; i < hi; ++i) {
        const int c = a[i];
        a[i] = run;
        run += c;
    }
    if (t == BS - 1) a[n] = before + incl;
    lds_barrier();
}

// ---------------------------------------------------------------- thermal sums
//
// The reference's thermal energy and pressure take every velocity relative to
// the centre-of-mass velocity, v - v_com, before squaring (energy.rs:25-37,
// pressure.rs:11-14).  A one-pass sum of m |v|^2 minus M |v_com|^2 would
// cancel catastrophically when the centre of mass drifts fast against the
// thermal speed.  The per-block partials therefore sum m |v - v_ref|^2 about a
// reference velocity that is already close to v_com: the centre-of-mass
// velocity of the group sums stored by the previous reduction.  The finalize
// then moves the reduced sum to the new centre of mass exactly,
//   sum m |v - c|^2 = sum m |v - v_ref|^2 - M |c - v_ref|^2,
// which loses nothing while v_ref ~ c, and stores it: stored group sums hold
// the thermal sums (about their own centre of mass) in S_MV2 / S_KV2.

// Component k of the reference velocity that partials of group sums S
// (stored by the last reduction; M the group's mass) are taken about.
__host__ __device__ __forceinline__ double sums_vref(const double* S, double M, int k) {
    const double v = M > 0.0 ? S[S_MVX + k] / M : 0.0;
    return isfinite(v) ? v : 0.0;
}

// sum m |v - v_ref|^2 moved to the centre of mass c = mv / M.
__host__ __device__ __forceinline__ double sums_about_com(double q, const double* mv, double M, const double* vref) {
    if (!(M > 0.0)) return q;
    const double dx = mv[0] / M - vref[0], dy = mv[1] / M - vref[1], dz = mv[2] / M - vref[2];
    return q - M * ((dx * dx + dy * dy) + dz * dz);
}

// ---------------------------------------------------------------- pair terms

// potential.rs:201-211 with LJ of potential.rs:57-69, FP64, reference order.
// potential.rs:57-75 and :206-211 for one pair: the terms added to the
// force, potential and virial accumulators (in this order, by the caller)
struct PairTerms64 {
    double gx, gy, gz, u, t;
};
__device__ __forceinline__ PairTerms64 pair_exact_terms(double dx, double dy, double dz, double r2,
                                                        const PairParam& p) {
    const double r = sqrt(r2);
    const double sr = p.sigma / r;
    const double s2 = sr * sr;
    const double s4 = s2 * s2;
    const double s6 = s2 * s4;
    const double s12 = s6 * s6;
    PairTerms64 q;
    q.u = 4.0 * p.eps * (s12 - s6) - p.u_cut;
    const double du = (24.0 * p.eps / r) * (s6 - 2.0 * s12);
    q.gx = (dx / r) * du;
    q.gy = (dy / r) * du;
    q.gz = (dz / r) * du;
    q.t = (q.gx * dx + q.gy * dy) + q.gz * dz;
    return q;
}

__device__ __forceinline__ void pair_exact(double dx, double dy, double dz, double r2,
                                           const PairParam& p, double& fx, double& fy,
                                           double& fz, double& pot, double& vir) {
    const PairTerms64 q = pair_exact_terms(dx, dy, dz, r2, p);
    fx += q.gx;
    fy += q.gy;
    fz += q.gz;
    pot += q.u;
    vir += q.t;
}

// Same physics in FP32 from the FP64 separation: F/r = 24 eps (s6 - 2 s12) / r^2,
// accumulated with fused multiply-adds: the per-pair values and the update.
struct PairTerms32 {
    float ax, ay, az, fr, r2, pu;
};
__device__ __forceinline__ PairTerms32 pair_fp32_terms(double dx, double dy, double dz, const PairParam& p) {
    PairTerms32 q;
    q.ax = (float)dx;
    q.ay = (float)dy;
    q.az = (float)dz;
    q.r2 = fmaf(q.ax, q.ax, fmaf(q.ay, q.ay, q.az * q.az));
    const float ir2 = 1.0f / q.r2;
    const float s2 = p.sig2f * ir2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    q.fr = p.eps24f * ir2 * fmaf(-2.0f, s12, s6);
    q.pu = fmaf(p.eps4f, s12 - s6, -p.ucutf);
    return q;
}

__device__ __forceinline__ void pair_fp32_add(const PairTerms32& q, float& fx, float& fy, float& fz, float& pot,
                                              float& vir) {
    fx = fmaf(q.ax, q.fr, fx);
    fy = fmaf(q.ay, q.fr, fy);
    fz = fmaf(q.az, q.fr, fz);
    pot += q.pu;
    vir = fmaf(q.fr, q.r2, vir);
}

__device__ __forceinline__ void pair_fp32(double dx, double dy, double dz, const PairParam& p,
                                          float& fx, float& fy, float& fz, float& pot,
                                          float& vir) {
    pair_fp32_add(pair_fp32_terms(dx, dy, dz, p), fx, fy, fz, pot, vir);
}

template <typename FT>
struct Acc;
template <>
struct Acc<double> {
    typedef PairTerms64 Terms;
    double fx = 0.0, fy = 0.0, fz = 0.0, pot = 0.0, vir = 0.0;
    __device__ __forceinline__ void add(double dx, double dy, double dz, double r2,
                                        const PairParam& p) {
        pair_exact(dx, dy, dz, r2, p, fx, fy, fz, pot, vir);
    }
    static __device__ __forceinline__ Terms terms(double dx, double dy, double dz, double r2, const PairParam& p) {
        return pair_exact_terms(dx, dy, dz, r2, p);
    }
    __device__ __forceinline__ void add_terms(const Terms& q) {
        fx += q.gx;
        fy += q.gy;
        fz += q.gz;
        pot += q.u;
        vir += q.t;
    }
};
template <>
struct Acc<float> {
    typedef PairTerms32 Terms;
    float fx = 0.0f, fy = 0.0f, fz = 0.0f, pot = 0.0f, vir = 0.0f;
    __device__ __forceinline__ void add(double dx, double dy, double dz, double,
                                        const PairParam& p) {
        pair_fp32(dx, dy, dz, p, fx, fy, fz, pot, vir);
    }
    static __device__ __forceinline__ Terms terms(double dx, double dy, double dz, double, const PairParam& p) {
        return pair_fp32_terms(dx, dy, dz, p);
    }
    __device__ __forceinline__ void add_terms(const Terms& q) { pair_fp32_add(q, fx, fy, fz, pot, vir); }
};

// Inclusive scan of one int per lane over a wave (the DPP tree of wave_total).
__device__ __forceinline__ int wave_scan_incl(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Two in-place exclusive scans of a[0..n) and b[0..n) (n <= kBrickCells) with
// shared barriers; a[n] and b[n] receive the totals.
template <int BS>
__device__ void block_exclusive_scan2(int* a, int* b, int n) {
    constexpr int NW = BS / 64;
    __shared__ int wsa[NW], wsb[NW];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int chunk = (n + BS - 1) / BS;
    const int lo = t * chunk < n ? t * chunk : n;
    const int hi = lo + chunk < n ? lo + chunk : n;
    int la = 0, lb = 0;
    for (int i = lo; i < hi; ++i) {
        la += a[i];
        lb += b[i];
    }
    const int ia = wave_scan_incl(la), ib = wave_scan_incl(lb);
    if (lane == 63) {
        wsa[w] = ia;
        wsb[w] = ib;
    }
    lds_barrier();
    int pa = 0, pb = 0;
    for (int i = 0; i < w; ++i) {
        pa += wsa[i];
        pb += wsb[i];
    }
    int ra = pa + ia - la, rb = pb + ib - lb;
    for (int i = lo; i < hi; ++i) {
        const int ca = a[i], cb = b[i];
        a[i] = ra;
        b[i] = rb;
        ra += ca;
        rb += cb;
    }
    if (t == BS - 1) {
        a[n] = pa + ia;
        b[n] = pb + ib;
    }
    lds_barrier();
}

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
        d.cell_slots[buf][(int64_t)k * d.cell_plane + cell] = (int32_t)slot;
        // within a quarter of the capacity: report, the host grows it between chunks
        if (4 * ((int)k + 1) > 3 * d.cell_cap) atomicMax(&d.stats->max_cell, (int)k + 1);
    } else {
        atomicOr(&d.stats->cell_overflow, 1u);
        atomicMax(&d.stats->max_cell, (int)k + 1);
    }
}

// Per-launch kernel timestamps: first block start to last block end, in
// wall-clock ticks, accumulated per launch with atomics (only when profiling).
__device__ __forceinline__ uint64_t prof_now(const DevState& d) {
    return d.prof ? (uint64_t)wall_clock64() : 0ull;
}

__device__ __forceinline__ void prof_mark(const DevState& d, int64_t launch, int kernel, uint64_t t0) {
    if (d.prof && threadIdx.x == 0) {
        const uint64_t t1 = (uint64_t)wall_clock64();
        unsigned long long* e = d.prof + ((launch % kProfRing) * P_NUM + kernel) * 2;
        atomicMin(e, (unsigned long long)t0);
        atomicMax(e + 1, (unsigned long long)t1);
    }
}

// Start and end of a whole block of k_mlp_net, merged with atomics into the
// first two words of its per-block area (every thread of the block calls it:
// it holds a barrier, and d.prof is the same for all).
__device__ __forceinline__ void prof_span_atomic(const DevState& d, int slot, uint64_t t0) {
    if (!d.prof) return;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long* e = d.prof_blk + (int64_t)slot * d.n_blocks * 3;
        atomicMin(e, (unsigned long long)t0);
        atomicMax(e + 1, (unsigned long long)wall_clock64());
    }
}

// Per-block record of a per-slot kernel (thread 0, after the block's last
// barrier); blocks that did no work store an empty interval.
__device__ __forceinline__ void prof_block(const DevState& d, int slot, uint64_t t0, bool worked,
                                           uint64_t entries) {
    if (d.prof && threadIdx.x == 0) {
        unsigned long long* e = d.prof_blk + ((int64_t)slot * d.n_blocks + blockIdx.x) * 3;
        e[0] = worked ? (unsigned long long)t0 : ~0ull;
        e[1] = worked ? (unsigned long long)wall_clock64() : 0ull;
        e[2] = entries;
    }
}

}  // namespace md
