// md_noalloc.hip -- md_macro_noalloc: the macro parameters of one particle
// type (macro_parameters/*.rs) from state arrays the caller holds in device
// memory, for callers that cannot allocate (code under HIP graph capture,
// tight serving loops).
//
// The entry point allocates nothing: its block partials live in `scratch`,
// a fixed array in this file's device data, one copy per GPU for the life of
// the process.  It launches four kernels on the caller's stream and neither
// synchronises nor copies to the host:
//   pass 1  per-block sums of m v, m |v|^2, potential and virial
//   reduce  fixed-order sum of the block rows; centre-of-mass velocity
//   pass 2  per-block sums of m |v - v_com|^2 (no cancellation)
//   final   fixed-order sum; KE, PE, thermal energy, T, P, momentum, v_com
// Every row a call reads was written earlier in the same call, so nothing
// needs clearing.  Calls that overlap on one GPU share its scratch and race.
#include <math.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "md_common.h"
#include "moldyn_hip.h"

void md_set_global_error(const std::string& msg);

namespace md {
namespace {

constexpr int kNaBS = 256;          // threads per block
constexpr int kNaBlocks = 1024;     // most blocks of a pass (rows of partials)
constexpr int kNaMaxDevices = 64;
constexpr int64_t kNaMaxAtoms = (int64_t)1 << 40;
// column k of block b at k * kNaBlocks + b; then the totals row (NS) and
// the centre-of-mass velocity (3)
constexpr int kNaTot = NS * kNaBlocks;
constexpr int kNaCom = kNaTot + NS;
constexpr int kNaWords = kNaCom + 4;

constexpr uint32_t kNaPass1 = (1u << S_MVX) | (1u << S_MVY) | (1u << S_MVZ) | (1u << S_MV2) | (1u << S_POT) |
                              (1u << S_VIR);
constexpr uint32_t kNaPass2 = 1u << S_KV2;  // sum m |v - v_com|^2

}  // namespace
}  // namespace md

// 106 KB per GPU
static __device__ __align__(256) double scratch[md::kNaWords];

namespace md {
namespace {

// Grid-stride pass over the particles (v: n x 3, pot / vir: n, may be null).
template <int PASS>
__global__ __launch_bounds__(kNaBS) void k_na_sums(const double* __restrict__ v, const double* __restrict__ pot,
                                                   const double* __restrict__ vir, int64_t n, double m,
                                                   double* __restrict__ scr) {
    double part[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) part[k] = 0.0;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    if (PASS == 2) {
        cx = scr[kNaCom];
        cy = scr[kNaCom + 1];
        cz = scr[kNaCom + 2];
    }
    const int64_t stride = (int64_t)gridDim.x * kNaBS;
    for (int64_t i = (int64_t)blockIdx.x * kNaBS + threadIdx.x; i < n; i += stride) {
        const double vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
        if (PASS == 1) {
            // mod.rs:17-24, mod.rs:30-32, energy.rs:14-22, energy.rs:40-49, pressure.rs:17
            part[S_MVX] += vx * m;
            part[S_MVY] += vy * m;
            part[S_MVZ] += vz * m;
            part[S_MV2] += m * ((vx * vx + vy * vy) + vz * vz);
            if (pot) part[S_POT] += pot[i];
            if (vir) part[S_VIR] += vir[i];
        } else {
            // energy.rs:25-37, pressure.rs:13-16: v - v_com before squaring
            const double dx = vx - cx, dy = vy - cy, dz = vz - cz;
            part[S_KV2] += m * ((dx * dx + dy * dy) + dz * dz);
        }
    }
    block_store<kNaBS>(part, scr + blockIdx.x, kNaBlocks, PASS == 1 ? kNaPass1 : kNaPass2);
}

// One block: the rows of this call's `nb` blocks in a fixed order.  Pass 1
// leaves the centre-of-mass velocity for pass 2 (mod.rs:12-25, with the
// caller's sequential fold of the masses); pass 2 writes the outputs.
template <int PASS>
__global__ __launch_bounds__(kNaBS) void k_na_reduce(int nb, int64_t n, double mass_sum, double volume,
                                                     double* __restrict__ out, double* __restrict__ scr) {
    constexpr uint32_t mask = PASS == 1 ? kNaPass1 : kNaPass2;
    double acc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kNaBS)
#pragma unroll
        for (int k = 0; k < NS; ++k)
            if ((mask >> k) & 1u) acc[k] += scr[k * kNaBlocks + b];
    block_store<kNaBS>(acc, scr + kNaTot, 1, mask);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double* tot = scr + kNaTot;
    if (PASS == 1) {
        scr[kNaCom] = tot[S_MVX] / mass_sum;
        scr[kNaCom + 1] = tot[S_MVY] / mass_sum;
        scr[kNaCom + 2] = tot[S_MVZ] / mass_sum;
        return;
    }
    const double th2 = tot[S_KV2];
    const double thermal = th2 / 2.0;
    out[0] = tot[S_MV2] / 2.0;                                        // energy.rs:14-22
    out[1] = tot[S_POT] / 2.0;                                        // energy.rs:40-49
    out[2] = thermal;                                                 // energy.rs:25-37
    out[3] = (2.0 * thermal) / (3.0 * (double)n * kKB) * 100.0;       // temperature.rs:4-7
    out[4] = (th2 + (-tot[S_VIR]) * 0.5) / volume / 3.0;              // pressure.rs:5-20
    out[5] = tot[S_MVX];                                              // mod.rs:28-34
    out[6] = tot[S_MVY];
    out[7] = tot[S_MVZ];
    out[8] = scr[kNaCom];                                             // mod.rs:12-25
    out[9] = scr[kNaCom + 1];
    out[10] = scr[kNaCom + 2];
}

// This GPU's scratch, looked up once per device (each GPU holds its own copy
// of the array at its own address; the host-side name is not that address).
double* g_scratch[kNaMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_on_current_device(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        md_set_global_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
        return MD_ERR_NO_DEVICE;
    }
    if (dev < 0 || dev >= kNaMaxDevices) {
        md_set_global_error("device index beyond the scratch table");
        return MD_ERR_NO_DEVICE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            md_set_global_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return MD_ERR_HIP;
        }
        g_scratch[dev] = (double*)p;
    }
    *out = g_scratch[dev];
    return MD_OK;
}

// macro_parameters/mod.rs:17-22: the sequential fold of 1.0 * m, cached for
// the last (m, n) a thread asked for
double mass_fold(double m, int64_t n) {
    thread_local double last_m = NAN, last_sum = 0.0;
    thread_local int64_t last_n = -1;
    if (n == last_n && m == last_m) return last_sum;
    double a = 0.0;
    for (int64_t i = 0; i < n; ++i) a = a + 1.0 * m;
    last_m = m;
    last_n = n;
    last_sum = a;
    return a;
}

}  // namespace
}  // namespace md

extern "C" int md_macro_noalloc(const double* v, const double* pot, const double* virial, int64_t n, double mass,
                                const double box[3], double* out, void* stream) {
    using namespace md;
    if (!v || !out || !box) {
        md_set_global_error("md_macro_noalloc: NULL argument");
        return MD_ERR_ARG;
    }
    if (n < 1 || n > kNaMaxAtoms) {
        md_set_global_error("md_macro_noalloc: particle count must be in [1, 2^40]");
        return MD_ERR_ARG;
    }
    if (!(mass > 0.0)) {
        md_set_global_error("md_macro_noalloc: mass must be positive");
        return MD_ERR_ARG;
    }
    double* scr = nullptr;
    int rc = scratch_on_current_device(&scr);
    if (rc) return rc;
    const int64_t want = (n + kNaBS - 1) / kNaBS;
    const int nb = (int)(want < kNaBlocks ? want : kNaBlocks);
    const double volume = box[0] * box[1] * box[2];
    const double msum = mass_fold(mass, n);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_na_sums<1>, dim3(nb), dim3(kNaBS), 0, st, v, pot, virial, n, mass, scr);
    hipLaunchKernelGGL(k_na_reduce<1>, dim3(1), dim3(kNaBS), 0, st, nb, n, msum, volume, out, scr);
    hipLaunchKernelGGL(k_na_sums<2>, dim3(nb), dim3(kNaBS), 0, st, v, pot, virial, n, mass, scr);
    hipLaunchKernelGGL(k_na_reduce<2>, dim3(1), dim3(kNaBS), 0, st, nb, n, msum, volume, out, scr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        md_set_global_error(std::string("md_macro_noalloc launch: ") + hipGetErrorString(e));
        return MD_ERR_HIP;
    }
    return MD_OK;
}
