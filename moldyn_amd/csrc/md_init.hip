// md_init.hip -- the reference's initialisers on the device: a blank state
// (initialize_particles, solver/src/initializer/position.rs:116-137), the
// U / FCC lattices (position.rs:24-104) and Maxwell-Boltzmann velocities
// with pair negation (velocity.rs:6-29), written straight into the records.
#include "md_common.h"

namespace md {

// every slot of both record buffers: position 0 (NaN on padding slots),
// velocity and force 0, the group's type (initialize_particles)
template <typename FT>
__global__ __launch_bounds__(256) void k_init_blank(Rec<FT>* __restrict__ r0, Rec<FT>* __restrict__ r1,
                                                    const GroupInfo* __restrict__ groups, int n_groups,
                                                    int64_t n_slots) {
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n_slots) return;
    int g = 0;
    while (g + 1 < n_groups && s >= groups[g + 1].start) ++g;
    const GroupInfo& G = groups[g];
    const double p = s - G.start < G.count ? 0.0 : __builtin_nan("");
    Rec<FT> r{};
    r.x = r.y = r.z = p;
    r.vx = r.vy = r.vz = 0.0;
    r.fx = r.fy = r.fz = (FT)0;
    r.type = (uint32_t)G.type;
    r0[s] = r;
    r1[s] = r;
}

// UnitCell::{U, FCC}: site k of cell (x, y, z), cell = x gy gz + y gz + z, at
// start + (X + o_k) a with o_0 = 0 and the FCC face centres o_1..o_3
template <typename FT>
__global__ __launch_bounds__(256) void k_init_lattice(Rec<FT>* __restrict__ rec, int64_t slot0, int64_t sites,
                                                      int fcc, double s0, double s1, double s2, int64_t gy,
                                                      int64_t gz, double a) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= sites) return;
    const int64_t cell = fcc ? i / 4 : i;
    const int k = fcc ? (int)(i % 4) : 0;
    const double X = (double)(cell / (gy * gz)), Y = (double)((cell / gz) % gy), Z = (double)(cell % gz);
    const double ox = (k == 2 || k == 3) ? 0.5 : 0.0;
    const double oy = (k == 1 || k == 3) ? 0.5 : 0.0;
    const double oz = (k == 1 || k == 2) ? 0.5 : 0.0;
    Rec<FT>& r = rec[slot0 + i];
    r.x = s0 + (X + ox) * a;
    r.y = s1 + (Y + oy) * a;
    r.z = s2 + (Z + oz) * a;
}

// SplitMix64 output number k (0-based) of the stream seeded with `seed`
__device__ __forceinline__ uint64_t sm64_at(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double open01_at(uint64_t seed, uint64_t k) {
    return ((double)(sm64_at(seed, k) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// standard normal number j of the sequence: Box-Muller pair j / 2 from
// uniforms 2 (j / 2) and 2 (j / 2) + 1, cosine for even j, sine for odd j
__device__ __forceinline__ double normal_at(uint64_t seed, uint64_t j) {
    const uint64_t p = j >> 1;
    const double u1 = open01_at(seed, 2 * p), u2 = open01_at(seed, 2 * p + 1);
    const double r = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    return (j & 1) ? r * sin(th) : r * cos(th);
}

// particle i < n / 2 gets sigma (N_3i, N_3i+1, N_3i+2), particle i + n / 2
// its negation; an odd last particle keeps its velocity
template <typename FT>
__global__ __launch_bounds__(256) void k_init_mb(Rec<FT>* __restrict__ rec, int64_t slot0, int64_t n, double sigma,
                                                 uint64_t seed) {
    const int64_t h = n / 2;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= h) return;
    const double vx = sigma * normal_at(seed, 3 * (uint64_t)i);
    const double vy = sigma * normal_at(seed, 3 * (uint64_t)i + 1);
    const double vz = sigma * normal_at(seed, 3 * (uint64_t)i + 2);
    Rec<FT>& a = rec[slot0 + i];
    a.vx = vx;
    a.vy = vy;
    a.vz = vz;
    Rec<FT>& b = rec[slot0 + i + h];
    b.vx = -vx;
    b.vy = -vy;
    b.vz = -vz;
}

static inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_init_blank(const DevState& d, bool fp32, const GroupInfo* groups, int n_groups, int64_t n_slots,
                             hipStream_t st) {
    if (n_slots <= 0) return hipSuccess;
    if (fp32)
        hipLaunchKernelGGL(k_init_blank<float>, dim3(blocks_for(n_slots)), dim3(256), 0, st, (Rec<float>*)d.rec[0],
                           (Rec<float>*)d.rec[1], groups, n_groups, n_slots);
    else
        hipLaunchKernelGGL(k_init_blank<double>, dim3(blocks_for(n_slots)), dim3(256), 0, st, (Rec<double>*)d.rec[0],
                           (Rec<double>*)d.rec[1], groups, n_groups, n_slots);
    return hipGetLastError();
}

hipError_t launch_init_lattice(void* rec, bool fp32, int64_t slot0, int64_t sites, bool fcc, const double start[3],
                               const int64_t grid[3], double a, hipStream_t st) {
    if (sites <= 0) return hipSuccess;
    if (fp32)
        hipLaunchKernelGGL(k_init_lattice<float>, dim3(blocks_for(sites)), dim3(256), 0, st, (Rec<float>*)rec, slot0,
                           sites, fcc ? 1 : 0, start[0], start[1], start[2], grid[1], grid[2], a);
    else
        hipLaunchKernelGGL(k_init_lattice<double>, dim3(blocks_for(sites)), dim3(256), 0, st, (Rec<double>*)rec,
                           slot0, sites, fcc ? 1 : 0, start[0], start[1], start[2], grid[1], grid[2], a);
    return hipGetLastError();
}

hipError_t launch_init_mb(void* rec, bool fp32, int64_t slot0, int64_t n, double sigma, uint64_t seed,
                          hipStream_t st) {
    if (n / 2 <= 0) return hipSuccess;
    if (fp32)
        hipLaunchKernelGGL(k_init_mb<float>, dim3(blocks_for(n / 2)), dim3(256), 0, st, (Rec<float>*)rec, slot0, n,
                           sigma, seed);
    else
        hipLaunchKernelGGL(k_init_mb<double>, dim3(blocks_for(n / 2)), dim3(256), 0, st, (Rec<double>*)rec, slot0, n,
                           sigma, seed);
    return hipGetLastError();
}

}  // namespace md
