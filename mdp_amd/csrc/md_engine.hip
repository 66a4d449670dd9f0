// Device-resident SNAP molecular dynamics (mdp_md_* in mdpsnap.h).
//
// Per step (velocity Verlet, NVE, metal units; src/C++/Core/cpuVelocityIntegration.cpp:307-340 and
// src/C++/Integration/Integration.cpp:17-156 for the upstream loop):
//   1. v += dt/2 f/m ; x += dt v
//   2. every `rebuild_every` steps: wrap x into the box, bin atoms into cells, build the full
//      neighbour list at rcut+skin (a workgroup per cell stages the 27 surrounding cells in LDS)
//      and its transpose (incoming entries per atom, sorted)
//   3. pair list: neighbour entries with d^2 < rc(ti,tj)^2 and d^2 > 1e-20 (cpuSnap.cpp:258-286),
//      compacted in CSR order by centre atom
//   4. SNAP: Utot (k_ui), Y + energy partials (k_yi), dE/dr per pair (k_dedr)
//   5. tally per atom without atomics: f_i = sum over own pairs - sum over incoming pairs,
//      W_i = -sum_j r_ij (x) dE/dr_ij, E_i from the energy partials; then v += dt/2 f/m.
// Neighbour entries are (owner j, image shift) pairs stored slot-major (slot k of atom i at k*N + i,
// so that per-atom loops read coalesced across the lanes of a wave); positions stay unwrapped between rebuilds and
// the displacement of a neighbour is always (x_j + shift*L) - x_i, evaluated without contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/mdpsnap.h"
#include "mlp_launch.hpp"
#include "runtime.hpp"
#include "snap_device.hpp"
#include "zbl.hpp"

using namespace mdp;

namespace {

constexpr int NBMAX = 96;      // neighbour slots per atom at rcut+skin
constexpr int CELLCAP = 48;    // atoms per cell accepted (bounds the per-cell insertion sort)
constexpr int SHK = 8;         // image shifts are coded in [-SHK, SHK] per axis
constexpr double FTM2V = 1.0 / 1.0364269e-4;   // metal units, eV/A/amu -> A/ps^2
constexpr double MVV2E = 1.0364269e-4;
constexpr double BOLTZ = 8.617343e-5;                // metal units, eV/K

enum Timer { T_VV1, T_REBUILD, T_PAIRS, T_UI, T_YI, T_DEDR, T_TALLY, T_SUR, T_AUDIT, T_PAIRPOT, T_COUNT };
const char* const kTimerNames[T_COUNT] = {"integrate", "neighbor_build", "pair_list", "compute_ui",
                                           "compute_yi", "compute_dedr", "tally_integrate", "surrogate_force",
                                           "audit", "pair_potential"};

__host__ __device__ inline int shift_code(int sx, int sy, int sz) {
    const int w = 2 * SHK + 1;
    return (sx + SHK) + w * ((sy + SHK) + w * (sz + SHK));
}
__host__ __device__ inline void shift_decode(int c, int& sx, int& sy, int& sz) {
    const int w = 2 * SHK + 1;
    sx = c % w - SHK;
    sy = (c / w) % w - SHK;
    sz = c / (w * w) - SHK;
}

// squared distance and image displacement evaluated without FMA contraction, so that neighbour and
// pair membership are reproducible bit for bit on the host
__device__ __forceinline__ double dist2_nc(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}
__device__ __forceinline__ double image_nc(double xj, int s, double L, double xi) {
#pragma clang fp contract(off)
    return (xj + s * L) - xi;
}

// ---- kernels -------------------------------------------------------------------------------------

__global__ void k_vv1(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ f,
                      const int* __restrict__ type, const double* __restrict__ massinv_t, double dtf, double dt, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double dtfm = dtf * massinv_t[type[i]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double vi = v[3 * i + d] + dtfm * f[3 * i + d];
        v[3 * i + d] = vi;
        x[3 * i + d] += dt * vi;
    }
}

// ---- Nose-Hoover chain thermostat (cpuVelocityIntegration.cpp:505-648) ----------------------------
// Chain state in one device array: eta_mass, eta, eta_dot (one extra zero entry), eta_dotdot, then the
// temperature, the velocity factor of the last thermostat half-step and the target temperature.
constexpr int NH_MAX = 10;
constexpr int NH_MASS = 0, NH_ETA = NH_MAX, NH_DOT = 2 * NH_MAX, NH_DDOT = 3 * NH_MAX + 1, NH_T = 4 * NH_MAX + 1,
              NH_FACTOR = NH_T + 1, NH_TTARGET = NH_T + 2, NH_SIZE = NH_T + 3;

// block sums of m v^2 (cpuComputeTempScalar, cpuComputes.cpp:276-296), fixed-order tree per block
__global__ __launch_bounds__(256) void k_ke_part(const double* __restrict__ v, const int* __restrict__ type,
                                                 const double* __restrict__ mass_t, int N, double* __restrict__ part) {
    __shared__ double red[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double s = 0.0;
    if (i < N) s = (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]) * mass_t[type[i]];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// One thermostat half-step of the chain (one thread): the temperature from the block sums, then the
// chain update of cpuNoseHooverThermostat with its nc_tchain sub-steps.  The velocities are scaled by
// the product of the sub-step factors, written to nh[NH_FACTOR] for the kernel that applies it.
__global__ void k_nh_chain(const double* __restrict__ part, int nparts, double* __restrict__ nh, double t_target,
                           double t_freq, double tdof, double boltz, double drag, double dt, double tfactor,
                           int mtchain, int nc_tchain, int eta_mass_flag) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double ke = 0.0;
    for (int b = 0; b < nparts; ++b) ke += part[b];
    double* eta_mass = nh + NH_MASS;
    double* eta = nh + NH_ETA;
    double* eta_dot = nh + NH_DOT;
    double* eta_dotdot = nh + NH_DDOT;
    double t_current = ke * tfactor;
    const double dthalf = 0.5 * dt, dt4 = 0.25 * dt, dt8 = 0.125 * dt;
    const double tdrag_factor = 1.0 - (dt * t_freq * drag / nc_tchain);
    double kecurrent = tdof * boltz * t_current;
    const double ke_target = tdof * boltz * t_target;
    if (eta_mass_flag) {
        eta_mass[0] = tdof * boltz * t_target / (t_freq * t_freq);
        for (int ich = 1; ich < mtchain; ++ich) eta_mass[ich] = boltz * t_target / (t_freq * t_freq);
    }
    eta_dotdot[0] = eta_mass[0] > 0.0 ? (kecurrent - ke_target) / eta_mass[0] : 0.0;
    const double ncfac = 1.0 / nc_tchain;
    double factor = 1.0;
    for (int iloop = 0; iloop < nc_tchain; ++iloop) {
        for (int ich = mtchain - 1; ich > 0; --ich) {
            const double expfac = exp(-ncfac * dt8 * eta_dot[ich + 1]);
            eta_dot[ich] *= expfac;
            eta_dot[ich] += eta_dotdot[ich] * ncfac * dt4;
            eta_dot[ich] *= tdrag_factor;
            eta_dot[ich] *= expfac;
        }
        double expfac = exp(-ncfac * dt8 * eta_dot[1]);
        eta_dot[0] *= expfac;
        eta_dot[0] += eta_dotdot[0] * ncfac * dt4;
        eta_dot[0] *= tdrag_factor;
        eta_dot[0] *= expfac;
        const double factor_eta = exp(-ncfac * dthalf * eta_dot[0]);
        factor *= factor_eta;
        t_current *= factor_eta * factor_eta;
        kecurrent = tdof * boltz * t_current;
        eta_dotdot[0] = eta_mass[0] > 0.0 ? (kecurrent - ke_target) / eta_mass[0] : 0.0;
        for (int ich = 0; ich < mtchain; ++ich) eta[ich] += ncfac * dthalf * eta_dot[ich];
        eta_dot[0] *= expfac;
        eta_dot[0] += eta_dotdot[0] * ncfac * dt4;
        eta_dot[0] *= expfac;
        for (int ich = 1; ich < mtchain; ++ich) {
            expfac = exp(-ncfac * dt8 * eta_dot[ich + 1]);
            eta_dot[ich] *= expfac;
            eta_dotdot[ich] = (eta_mass[ich - 1] * eta_dot[ich - 1] * eta_dot[ich - 1] - boltz * t_target) / eta_mass[ich];
            eta_dot[ich] += eta_dotdot[ich] * ncfac * dt4;
            eta_dot[ich] *= expfac;
        }
    }
    nh[NH_T] = t_current;
    nh[NH_FACTOR] = factor;
    nh[NH_TTARGET] = t_target;
}

// first half-step after the thermostat: v = v * factor + dtf/m f, x += dt v
__global__ void k_vv1_nh(double* __restrict__ x, double* __restrict__ v, const double* __restrict__ f,
                         const int* __restrict__ type, const double* __restrict__ massinv_t,
                         const double* __restrict__ nh, double dtf, double dt, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double fac = nh[NH_FACTOR];
    const double dtfm = dtf * massinv_t[type[i]];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double vi = v[3 * i + d] * fac + dtfm * f[3 * i + d];
        v[3 * i + d] = vi;
        x[3 * i + d] += dt * vi;
    }
}

__global__ void k_vscale_nh(double* __restrict__ v, const double* __restrict__ nh, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * N) return;
    v[i] *= nh[NH_FACTOR];
}

// block sums for the thermo output: [0] energy, [1..3] virial xx, yy, zz, [4] m v^2 (fixed-order tree)
__global__ __launch_bounds__(256) void k_thermo_part(const double* __restrict__ eatom, const double* __restrict__ vir,
                                                     const double* __restrict__ v, const int* __restrict__ type,
                                                     const double* __restrict__ mass_t, int N,
                                                     double* __restrict__ part) {
    __shared__ double red[5][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    double q[5] = {0, 0, 0, 0, 0};
    if (i < N) {
        q[0] = eatom[i];
        q[1] = vir[6 * i];
        q[2] = vir[6 * i + 1];
        q[3] = vir[6 * i + 2];
        q[4] = (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]) * mass_t[type[i]];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = q[k];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w)
#pragma unroll
            for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 5) part[5 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// wrap into [0, L), record the maximum squared displacement since the last build, bin into cells
__global__ void k_wrap_bin(double* __restrict__ x, double* __restrict__ xref, int* __restrict__ cell_of,
                           int* __restrict__ cell_cnt, unsigned long long* __restrict__ maxd2_bits, double Lx,
                           double Ly, double Lz, int ncx, int ncy, int ncz, int N, int check,
                           int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double L[3] = {Lx, Ly, Lz};
    const int nc[3] = {ncx, ncy, ncz};
    double d2 = 0.0;
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double xi = x[3 * i + d];
        if (!isfinite(xi)) {   // keep the cell index in range; the host reports the error
            atomicOr(flags, 8);
            xi = 0.0;
        }
        if (check) {
            const double dd = xi - xref[3 * i + d];
            d2 += dd * dd;
        }
        xi -= floor(xi / L[d]) * L[d];
        if (xi >= L[d]) xi -= L[d];
        if (xi < 0.0) xi = 0.0;
        if (!(xi < L[d])) {   // a runaway coordinate (|x| beyond 2^53 box lengths) does not wrap: report it
            atomicOr(flags, 8);
            xi = 0.0;
        }
        x[3 * i + d] = xi;
        xref[3 * i + d] = xi;
        c[d] = min(static_cast<int>(xi * nc[d] / L[d]), nc[d] - 1);
    }
    if (check) atomicMax(maxd2_bits, static_cast<unsigned long long>(__double_as_longlong(d2)));
    const int cid = c[0] + ncx * (c[1] + ncy * c[2]);
    cell_of[i] = cid;
    atomicAdd(&cell_cnt[cid], 1);
}

__global__ void k_cell_fill(const int* __restrict__ cell_of, const int* __restrict__ cell_start,
                            int* __restrict__ cell_fill, int* __restrict__ cell_atoms, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int c = cell_of[i];
    const int pos = atomicAdd(&cell_fill[c], 1);
    cell_atoms[cell_start[c] + pos] = i;
}

// ascending atom order inside each cell (deterministic list order)
__global__ void k_cell_sort(const int* __restrict__ cell_start, int* __restrict__ cell_atoms, int ncell,
                            int* __restrict__ overflow) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    const int b = cell_start[c], e = cell_start[c + 1];
    if (e - b > CELLCAP) atomicOr(overflow, 2);
    for (int k = b + 1; k < e; ++k) {
        const int v = cell_atoms[k];
        int m = k - 1;
        while (m >= b && cell_atoms[m] > v) {
            cell_atoms[m + 1] = cell_atoms[m];
            --m;
        }
        cell_atoms[m + 1] = v;
    }
}

// Eight lanes per centre atom, atoms taken in cell order (so the atoms of a wave share their 27
// surrounding cells in L1).  The group walks the 27 cells (periodic images applied) in a fixed order and
// each cell's atoms in ascending order, eight candidates per pass; accepted candidates are ranked by a
// ballot, which keeps the list in candidate order.
constexpr int NSUB = 8;

__global__ __launch_bounds__(256) void k_neighbors_cell(const double* __restrict__ x, const int* __restrict__ cell_of,
                                                        const int* __restrict__ cell_start,
                                                        const int* __restrict__ cell_atoms, int* __restrict__ nb_cnt,
                                                        int* __restrict__ nb_j, int* __restrict__ nb_s, double Lx,
                                                        double Ly, double Lz, int ncx, int ncy, int ncz, double rcs2,
                                                        int N, int* __restrict__ overflow) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = t / NSUB, sub = t % NSUB, lane = threadIdx.x & 63;
    const unsigned long long group = 0xFFull << (lane & ~(NSUB - 1));
    const unsigned long long below = (1ull << lane) - 1;
    const bool live = g < N;
    const int i = live ? cell_atoms[g] : 0;
    const int c = live ? cell_of[i] : 0;
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    const double xi = live ? x[3 * i] : 0.0, yi = live ? x[3 * i + 1] : 0.0, zi = live ? x[3 * i + 2] : 0.0;
    const int self_code = shift_code(0, 0, 0);
    int cnt = 0;
    for (int k = 0; k < 27; ++k) {
        const int ux = cx + k % 3 - 1, uy = cy + (k / 3) % 3 - 1, uz = cz + k / 9 - 1;
        const int sx = (ux < 0) ? -1 : (ux >= ncx ? 1 : 0);
        const int sy = (uy < 0) ? -1 : (uy >= ncy ? 1 : 0);
        const int sz = (uz < 0) ? -1 : (uz >= ncz ? 1 : 0);
        const int nc = (ux - sx * ncx) + ncx * ((uy - sy * ncy) + ncy * (uz - sz * ncz));
        const int code = shift_code(sx, sy, sz);
        const int b = live ? cell_start[nc] : 0, n = live ? cell_start[nc + 1] - b : 0;
        for (int m0 = 0; __ballot(m0 < n) != 0; m0 += NSUB) {   // wave-uniform trip count
            const int m = m0 + sub;
            int j = 0;
            bool acc = false;
            if (m < n) {
                j = cell_atoms[b + m];
                const double d2 = dist2_nc(image_nc(x[3 * j], sx, Lx, xi), image_nc(x[3 * j + 1], sy, Ly, yi),
                                           image_nc(x[3 * j + 2], sz, Lz, zi));
                acc = d2 <= rcs2 && !(j == i && code == self_code);
            }
            const unsigned long long mk = __ballot(acc) & group;
            if (acc) {
                const int q = cnt + __popcll(mk & below);
                if (q < NBMAX) {
                    nb_j[static_cast<size_t>(q) * N + i] = j;
                    nb_s[static_cast<size_t>(q) * N + i] = code;
                }
            }
            cnt += __popcll(mk);
        }
    }
    if (live && sub == 0) {
        if (cnt > NBMAX) atomicOr(overflow, 1);
        nb_cnt[i] = min(cnt, NBMAX);
    }
}

// The same lists from a per-wavefront neighbour tile in LDS.  One wave owns one centre cell: it lists the
// 27 surrounding cells (periodic images applied) in the fixed order above, stages their atoms -- index,
// image shift and the imaged position x_j + s L -- in LDS, NTILE candidates at a time, and its centre atoms
// (eight lanes each, eight atoms per pass, ascending) scan the tile from LDS.  Candidates, their order and
// the distance arithmetic are those of k_neighbors_cell, so the lists are identical.
constexpr int NTILE = 256;

__global__ __launch_bounds__(64) void k_neighbors_tile(const double* __restrict__ x,
                                                       const int* __restrict__ cell_start,
                                                       const int* __restrict__ cell_atoms, int* __restrict__ nb_cnt,
                                                       int* __restrict__ nb_j, int* __restrict__ nb_s, double Lx,
                                                       double Ly, double Lz, int ncx, int ncy, int ncz, double rcs2,
                                                       int N, int* __restrict__ overflow) {
    __shared__ double tx[NTILE], ty[NTILE], tz[NTILE];
    __shared__ int tj[NTILE], ts[NTILE];
    __shared__ int k_off[28], k_base[27], k_code[27];
    __shared__ double k_dx[27], k_dy[27], k_dz[27];
    const int c = blockIdx.x, lane = threadIdx.x;
    const int cb = cell_start[c], cn = cell_start[c + 1] - cb;
    if (cn == 0) return;   // uniform over the workgroup
    const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
    if (lane < 27) {
        const int k = lane;
        const int ux = cx + k % 3 - 1, uy = cy + (k / 3) % 3 - 1, uz = cz + k / 9 - 1;
        const int sx = (ux < 0) ? -1 : (ux >= ncx ? 1 : 0);
        const int sy = (uy < 0) ? -1 : (uy >= ncy ? 1 : 0);
        const int sz = (uz < 0) ? -1 : (uz >= ncz ? 1 : 0);
        const int nc = (ux - sx * ncx) + ncx * ((uy - sy * ncy) + ncy * (uz - sz * ncz));
        k_base[k] = cell_start[nc];
        k_off[k + 1] = cell_start[nc + 1] - cell_start[nc];   // count, summed below
        k_code[k] = shift_code(sx, sy, sz);
        k_dx[k] = sx * Lx;
        k_dy[k] = sy * Ly;
        k_dz[k] = sz * Lz;
    }
    __syncthreads();
    if (lane == 0) {
        k_off[0] = 0;
        for (int k = 0; k < 27; ++k) k_off[k + 1] += k_off[k];
    }
    __syncthreads();
    const int total = k_off[27];
    const int g = lane / NSUB, sub = lane % NSUB;
    const unsigned long long group = 0xFFull << (lane & ~(NSUB - 1));
    const unsigned long long below = (1ull << lane) - 1;
    const int self_code = shift_code(0, 0, 0);
    for (int a0 = 0; a0 < cn; a0 += 64 / NSUB) {
        const bool live = a0 + g < cn;
        const int i = live ? cell_atoms[cb + a0 + g] : 0;
        const double xi = live ? x[3 * i] : 0.0, yi = live ? x[3 * i + 1] : 0.0, zi = live ? x[3 * i + 2] : 0.0;
        int cnt = 0;
        for (int t0 = 0; t0 < total; t0 += NTILE) {
            const int nt = min(NTILE, total - t0);
            __syncthreads();   // the previous pass has finished reading the tile
            for (int t = lane; t < nt; t += 64) {
                const int tt = t0 + t;
                int k = 0;   // the cell holding candidate tt: k_off[k] <= tt < k_off[k + 1]
#pragma unroll
                for (int step = 16; step > 0; step >>= 1)
                    if (k + step < 27 && k_off[k + step] <= tt) k += step;
                const int j = cell_atoms[k_base[k] + tt - k_off[k]];
                {
#pragma clang fp contract(off)
                    tx[t] = x[3 * j] + k_dx[k];
                    ty[t] = x[3 * j + 1] + k_dy[k];
                    tz[t] = x[3 * j + 2] + k_dz[k];
                }
                tj[t] = j;
                ts[t] = k_code[k];
            }
            __syncthreads();
            for (int m0 = 0; m0 < nt; m0 += NSUB) {   // uniform trip count
                const int m = m0 + sub;
                bool acc = false;
                int j = 0;
                if (live && m < nt) {
                    j = tj[m];
                    double d2;
                    {
#pragma clang fp contract(off)
                        d2 = dist2_nc(tx[m] - xi, ty[m] - yi, tz[m] - zi);
                    }
                    acc = d2 <= rcs2 && !(j == i && ts[m] == self_code);
                }
                const unsigned long long mk = __ballot(acc) & group;
                if (acc) {
                    const int q = cnt + __popcll(mk & below);
                    if (q < NBMAX) {
                        nb_j[static_cast<size_t>(q) * N + i] = j;
                        nb_s[static_cast<size_t>(q) * N + i] = ts[m];
                    }
                }
                cnt += __popcll(mk);
            }
        }
        if (live && sub == 0) {
            if (cnt > NBMAX) atomicOr(overflow, 1);
            nb_cnt[i] = min(cnt, NBMAX);
        }
    }
}

// Small boxes (fewer than three cells along an axis): every atom against every periodic image.
__global__ void k_neighbors_brute(const double* __restrict__ x, int* __restrict__ nb_cnt, int* __restrict__ nb_j,
                                  int* __restrict__ nb_s, double Lx, double Ly, double Lz, double rcs, double rcs2,
                                  int N, int* __restrict__ overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int rx = min(static_cast<int>(ceil(rcs / Lx)) + 1, SHK), ry = min(static_cast<int>(ceil(rcs / Ly)) + 1, SHK),
              rz = min(static_cast<int>(ceil(rcs / Lz)) + 1, SHK);
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    int cnt = 0;
    for (int sz = -rz; sz <= rz; ++sz)
        for (int sy = -ry; sy <= ry; ++sy)
            for (int sx = -rx; sx <= rx; ++sx)
                for (int j = 0; j < N; ++j) {
                    if (j == i && sx == 0 && sy == 0 && sz == 0) continue;
                    const double d2 = dist2_nc(image_nc(x[3 * j], sx, Lx, xi), image_nc(x[3 * j + 1], sy, Ly, yi),
                                               image_nc(x[3 * j + 2], sz, Lz, zi));
                    if (d2 <= rcs2) {
                        if (cnt < NBMAX) {
                            nb_j[static_cast<size_t>(cnt) * N + i] = j;
                            nb_s[static_cast<size_t>(cnt) * N + i] = shift_code(sx, sy, sz);
                        }
                        ++cnt;
                    }
                }
    if (cnt > NBMAX) atomicOr(overflow, 1);
    nb_cnt[i] = min(cnt, NBMAX);
}

// total of the per-atom entry counts: one atomic per workgroup
__global__ __launch_bounds__(256) void k_sum_counts(const int* __restrict__ cnt, int N,
                                                    unsigned long long* __restrict__ total) {
    __shared__ unsigned long long part[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = i < N ? static_cast<unsigned long long>(cnt[i]) : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total, part[0] + part[1] + part[2] + part[3]);
}

// incoming entries: for every entry e = (i -> j) append e to j's list, then sort each list.  The lists are
// atom-major (slot k of atom j at j*NBMAX + k): the tally's 16 lanes per atom then read one contiguous run.
__global__ void k_incoming(const int* __restrict__ nb_cnt, const int* __restrict__ nb_j, int* __restrict__ in_cnt,
                           int* __restrict__ in_e, int N, int* __restrict__ overflow) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int k = 0; k < nb_cnt[i]; ++k) {
        const int e = k * N + i;
        const int j = nb_j[e];
        const int slot = atomicAdd(&in_cnt[j], 1);
        if (slot < NBMAX)
            in_e[static_cast<size_t>(j) * NBMAX + slot] = e;
        else
            atomicOr(overflow, 4);
    }
}

__global__ void k_incoming_sort(int* __restrict__ in_cnt, int* __restrict__ in_e, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int n = min(in_cnt[i], NBMAX);
    in_cnt[i] = n;
    int* a = in_e + static_cast<size_t>(i) * NBMAX;   // slot k at a[k]
    const size_t st = 1;
    for (int k = 1; k < n; ++k) {
        const int v = a[k * st];
        int m = k - 1;
        while (m >= 0 && a[m * st] > v) {
            a[(m + 1) * st] = a[m * st];
            --m;
        }
        a[(m + 1) * st] = v;
    }
}

__device__ __forceinline__ void entry_disp(const double* __restrict__ x, int i, int j, int code, double Lx, double Ly,
                                           double Lz, double& dx, double& dy, double& dz) {
    int sx, sy, sz;
    shift_decode(code, sx, sy, sz);
    dx = image_nc(x[3 * j], sx, Lx, x[3 * i]);
    dy = image_nc(x[3 * j + 1], sy, Ly, x[3 * i + 1]);
    dz = image_nc(x[3 * j + 2], sz, Lz, x[3 * i + 2]);
}

// pairs per atom (entries within the pair cutoff); 16 lanes per atom as in k_pair_fill_md below
__global__ __launch_bounds__(256) void k_pair_count_md(const double* __restrict__ x, const int* __restrict__ type,
                                                       const int* __restrict__ nb_cnt, const int* __restrict__ nb_j,
                                                       const int* __restrict__ nb_s, const double* __restrict__ rc2_tt,
                                                       int ntypes, int* __restrict__ pcnt, double Lx, double Ly,
                                                       double Lz, int N) {
    constexpr int SUB = 16;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / SUB, sub = t % SUB, lane = threadIdx.x & 63;
    const unsigned long long group = 0xFFFFull << (lane & ~(SUB - 1));
    const bool live = i < N;
    const int ti = live ? type[i] : 0;
    const int n = live ? nb_cnt[i] : 0;
    int c = 0;
    for (int k0 = 0; __ballot(k0 < n) != 0; k0 += SUB) {   // wave-uniform trip count
        const int k = k0 + sub;
        bool acc = false;
        if (k < n) {
            const int e = k * N + i, j = nb_j[e];
            double dx, dy, dz;
            entry_disp(x, i, j, nb_s[e], Lx, Ly, Lz, dx, dy, dz);
            const double d2 = dist2_nc(dx, dy, dz);
            acc = d2 < rc2_tt[ti * (ntypes + 1) + type[j]] && d2 > 1e-20;
        }
        c += __popcll(__ballot(acc) & group);
    }
    if (live && sub == 0) pcnt[i] = c;
}

// 16 lanes per atom: each pass tests 16 consecutive entries of the atom, and the accepted ones get
// consecutive pair slots from a ballot, so a wave's stores cover a few contiguous pair ranges (one per
// atom) instead of 64 scattered slots.  Pair order within an atom is the entry order, as with one lane
// per atom.
constexpr int PSUB = 16;

__global__ __launch_bounds__(256) void k_pair_fill_md(const double* __restrict__ x, const int* __restrict__ type,
                                                      const int* __restrict__ nb_cnt, const int* __restrict__ nb_j,
                                                      const int* __restrict__ nb_s, const double* __restrict__ rc2_tt,
                                                      int ntypes, const int* __restrict__ poff, int* __restrict__ pi_,
                                                      int* __restrict__ pj, int* __restrict__ pti,
                                                      int* __restrict__ ptj, double* __restrict__ prij,
                                                      int* __restrict__ entry_pair, double Lx, double Ly, double Lz,
                                                      int N, SnapDev sd, double* __restrict__ pgeo) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / PSUB, sub = t % PSUB, lane = threadIdx.x & 63;
    const unsigned long long group = 0xFFFFull << (lane & ~(PSUB - 1));
    const unsigned long long below = (1ull << lane) - 1;
    const bool live = i < N;
    const int ti = live ? type[i] : 0;
    const int n = live ? nb_cnt[i] : 0;
    int p = live ? poff[i] : 0;
    for (int k0 = 0; __ballot(k0 < n) != 0; k0 += PSUB) {   // wave-uniform trip count
        const int k = k0 + sub;
        const bool valid = k < n;
        int j = 0, tjv = 0;
        double dx = 0.0, dy = 0.0, dz = 0.0;
        bool acc = false;
        const int e = k * N + i;
        if (valid) {
            j = nb_j[e];
            tjv = type[j];
            entry_disp(x, i, j, nb_s[e], Lx, Ly, Lz, dx, dy, dz);
            const double d2 = dist2_nc(dx, dy, dz);
            acc = d2 < rc2_tt[ti * (ntypes + 1) + tjv] && d2 > 1e-20;
        }
        const unsigned long long m = __ballot(acc) & group;
        if (valid) {
            if (acc) {
                const int q = p + __popcll(m & below);
                pi_[q] = i;
                pj[q] = j;
                pti[q] = ti;
                ptj[q] = tjv;
                prij[3 * static_cast<size_t>(q)] = dx;
                prij[3 * static_cast<size_t>(q) + 1] = dy;
                prij[3 * static_cast<size_t>(q) + 2] = dz;
                // Cayley-Klein parameters and weighted switching factor, read by compute_ui
                PairGeom g;
                pair_geometry<false>(dx, dy, dz, (sd.radelem[ti] + sd.radelem[tjv]) * sd.rcutfac, sd.rmin0, sd.rfac0,
                                     sd.switchflag, sd.wjelem[tjv], g);
                double* __restrict__ o = pgeo + 5 * static_cast<size_t>(q);
                o[0] = g.ar;
                o[1] = g.ai;
                o[2] = g.br;
                o[3] = g.bi;
                o[4] = g.sfac;
                entry_pair[e] = q;
            } else {
                entry_pair[e] = -1;
            }
        }
        p += __popcll(m);
    }
}

// f, per-atom virial and energy of atom i without atomics; optionally the second velocity half-step.
// 16 lanes per atom: lane s takes the atom's pairs s, s+16, ... and incoming entries s, s+16, ... (the
// pair reads of a wave are then contiguous), and the 16 partial sums are combined by a butterfly.
__global__ __launch_bounds__(256) void k_tally(const double* __restrict__ dedr, const double* __restrict__ prij,
                                               const int* __restrict__ poff, const int* __restrict__ in_cnt,
                                               const int* __restrict__ in_e, const int* __restrict__ entry_pair,
                                               const double* __restrict__ epart, int nrows, double e0,
                                               const double* __restrict__ eext, double* __restrict__ f,
                                               double* __restrict__ vir, double* __restrict__ eatom,
                                               double* __restrict__ v, const int* __restrict__ type,
                                               const double* __restrict__ massinv_t, double dtf, int integrate, int N,
                                               unsigned long long* __restrict__ pair_sum,
                                               const double* __restrict__ zacc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / PSUB, sub = t % PSUB;
    const bool live = i < N;
    if (pair_sum && t == 0) atomicAdd(pair_sum, static_cast<unsigned long long>(poff[N]));   // pairs per step
    double a[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // f x, y, z; virial 0..5; energy
    if (live) {
        for (int p = poff[i] + sub; p < poff[i + 1]; p += PSUB) {
            const double gx = dedr[3 * static_cast<size_t>(p)], gy = dedr[3 * static_cast<size_t>(p) + 1],
                         gz = dedr[3 * static_cast<size_t>(p) + 2];
            const double rx = prij[3 * static_cast<size_t>(p)], ry = prij[3 * static_cast<size_t>(p) + 1],
                         rz = prij[3 * static_cast<size_t>(p) + 2];
            a[0] += gx;
            a[1] += gy;
            a[2] += gz;
            a[3] -= rx * gx;
            a[4] -= ry * gy;
            a[5] -= rz * gz;
            a[6] -= rx * gy;
            a[7] -= rx * gz;
            a[8] -= ry * gz;
        }
        const int nin = in_cnt[i];
        for (int k = sub; k < nin; k += PSUB) {
            const int p = entry_pair[in_e[static_cast<size_t>(i) * NBMAX + k]];
            if (p >= 0) {
                a[0] -= dedr[3 * static_cast<size_t>(p)];
                a[1] -= dedr[3 * static_cast<size_t>(p) + 1];
                a[2] -= dedr[3 * static_cast<size_t>(p) + 2];
            }
        }
        if (!eext)
            for (int r = sub; r < nrows; r += PSUB) a[9] += epart[static_cast<size_t>(r) * N + i];
    }
#pragma unroll
    for (int w = PSUB / 2; w >= 1; w /= 2)
#pragma unroll
        for (int c = 0; c < 10; ++c) a[c] += __shfl_xor(a[c], w);
    if (!live || sub != 0) return;
    double ez = 0.0;
    if (zacc) {   // reference pair potential of the same atom (k_zbl_md)
#pragma unroll
        for (int c = 0; c < 9; ++c) a[c] += zacc[10 * static_cast<size_t>(i) + c];
        ez = zacc[10 * static_cast<size_t>(i) + 9];
    }
    f[3 * i] = a[0];
    f[3 * i + 1] = a[1];
    f[3 * i + 2] = a[2];
#pragma unroll
    for (int c = 0; c < 6; ++c) vir[6 * i + c] = a[3 + c];
    eatom[i] = (eext ? eext[i] : e0 + (2.0 / 3.0) * a[9]) + ez;
    if (integrate) {
        const double dtfm = dtf * massinv_t[type[i]];
        v[3 * i] += dtfm * a[0];
        v[3 * i + 1] += dtfm * a[1];
        v[3 * i + 2] += dtfm * a[2];
    }
}

// second velocity half-step on its own (steps whose forces are chosen after an audit)
__global__ void k_vv2(double* __restrict__ v, const double* __restrict__ f, const int* __restrict__ type,
                      const double* __restrict__ massinv_t, double dtf, int N) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double dtfm = dtf * massinv_t[type[i]];
    v[3 * i] += dtfm * f[3 * i];
    v[3 * i + 1] += dtfm * f[3 * i + 1];
    v[3 * i + 2] += dtfm * f[3 * i + 2];
}

// Reference pair potential (ZBL) of every atom over its full neighbour list, with the upstream
// full-list conventions (cpuFullNeighPairList's d^2 <= rc^2 test, src/C++/Core/cpuNeighCount.cpp:60-81;
// cpuFullNeighPairDecomposition and cpuVirialPairTally with factor -1/2, cpuPotentials.cpp:1587-1649 and
// the sign flip of :1837): e_i = sum_j u(r_ij)/2, F_i = sum_j du/dxij, W_i = -1/2 sum_j xij (x) du/dxij,
// xij = x_j - x_i.  Each atom owns its sums, so there are no atomics; 16 lanes per atom take every 16th
// entry and combine by a butterfly.  zacc[10 i + c]: force x, y, z, virial xx, yy, zz, xy, xz, yz, energy.
__global__ __launch_bounds__(256) void k_zbl_md(const double* __restrict__ x, const int* __restrict__ nb_cnt,
                                                const int* __restrict__ nb_j, const int* __restrict__ nb_s,
                                                double Lx, double Ly, double Lz, int N, ZblParams zp,
                                                double* __restrict__ zacc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / PSUB, sub = t % PSUB;
    const bool live = i < N;
    double a[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (live) {
        const int n = nb_cnt[i];
        for (int k = sub; k < n; k += PSUB) {
            const int e = k * N + i;
            double dx, dy, dz;
            entry_disp(x, i, nb_j[e], nb_s[e], Lx, Ly, Lz, dx, dy, dz);
            const double d2 = dist2_nc(dx, dy, dz);
            if (d2 <= zp.rout2) {
                const double r = sqrt(d2);
                double u, dudr;
                zbl_eval(zp, r, u, dudr);
                const double g = dudr / r, gx = g * dx, gy = g * dy, gz = g * dz;
                a[0] += gx;
                a[1] += gy;
                a[2] += gz;
                a[3] -= 0.5 * dx * gx;
                a[4] -= 0.5 * dy * gy;
                a[5] -= 0.5 * dz * gz;
                a[6] -= 0.5 * dx * gy;
                a[7] -= 0.5 * dx * gz;
                a[8] -= 0.5 * dy * gz;
                a[9] += 0.5 * u;
            }
        }
    }
#pragma unroll
    for (int w = PSUB / 2; w >= 1; w /= 2)
#pragma unroll
        for (int c = 0; c < 10; ++c) a[c] += __shfl_xor(a[c], w);
    if (!live || sub != 0) return;
#pragma unroll
    for (int c = 0; c < 10; ++c) zacc[10 * static_cast<size_t>(i) + c] = a[c];
}

// per-block partial sums of |a - b|^2
__global__ __launch_bounds__(256) void k_diff_sumsq(const double* __restrict__ a, const double* __restrict__ b, long n,
                                                    double* __restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long k = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; k < n; k += static_cast<long>(gridDim.x) * 256) {
        const double d = a[k] - b[k];
        s += d * d;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (static_cast<int>(threadIdx.x) < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

inline int nb(long n, int b) { return static_cast<int>((n + b - 1) / b); }

}  // namespace

// ---- engine --------------------------------------------------------------------------------------

struct mdp_md {
    int N = 0, ntypes = 0, J = 0;
    double L[3] = {0, 0, 0};
    double dt = 0, dtf = 0, skin = 0, rcmax = 0, rcnb = 0;   // rcnb: neighbour-list cutoff (skin included)
    // reference pair potential (mdp_md_set_zbl); off: SNAP alone
    bool zbl = false;
    ZblParams zp;
    double zbl_rout = 0;
    int rebuild_every = 10;
    int ncell[3] = {0, 0, 0};
    bool use_cells = false;
    double e0 = 0;
    int nrows = 0;
    hipStream_t st = nullptr;
    DevTables tab;
    std::vector<double> massinv_h, rc2_h;
    // device state
    DevBuf x, v, f, type, massinv, xref, cell_of, cell_cnt, cell_start, cell_fill, cell_atoms, scan_tmp, nb_cnt, nb_j,
        nb_s, in_cnt, in_e, pcnt, poff, pi, pj, pti, ptj, prij, entry_pair, Ur, Ui, Yr, Yi, epart, dedr, eatom, vir,
        beta_t, rc2, flags, maxd2, row_order, yt, pair_sum, pgeo, zacc;
    // Nose-Hoover NVT (mdp_md_set_nvt); off: NVE
    bool nvt = false;
    double nvt_tstart = 0, nvt_tstop = 0, nvt_tfreq = 0, nvt_drag = 0, nvt_tdof = 0;
    int nvt_mtchain = 3, nvt_nc = 1, nvt_massflag = 1;
    long nvt_begin = 0, nvt_end = 1, nvt_step = 0;
    DevBuf nh, kepart, mass_t;
    long pair_steps0 = 0;   // force evaluations counted in pair_sum since the last reset
    long stats[6] = {0, 0, 0, 0, 0, 0};
    bool forces_current = false;
    long force_evals = 0;   // force evaluations (steps and compute calls) whose pairs were counted
    std::vector<double> rad_h, wj_h;
    // learned surrogate (mdp_md_set_surrogate): forces from a per-atom MLP over low-order bispectrum
    // descriptors, audited every audit_every steps against plain SNAP
    bool sur = false;
    int sur_nin = 0, sur_ntrip = 0, audit_every = 0;
    double audit_tol = 0.0;
    DevTables tab_s;
    DevBuf sUr, sUi, sB, sE, sG, sBeta, sYr, sYi, sParams, sPacked, sAff, sJjb, sFac, sRowOrder, fP, vP, eP, ssq;
    long audits = 0, fallbacks = 0;
    double rmse_last = 0, rmse_max = 0, rmse_sum = 0;
    // autotuned launch variants (see autotune())
    int dedr_variant = 3, yi_waves = 8;
    int ui_variant = 0;   // compute_ui: 0 = one thread per atom and row group, G > 0 = LDS sums on G workgroups
    int nb_variant = 0;   // cell-list neighbour kernel (neighbors_launch)
    bool tuned = false;
    std::string tuning_json = "{}";
    // timing
    bool timing = false;
    unsigned timing_mask = ~0u;   // timer classes recorded while timing is on
    std::vector<hipEvent_t> ev_pool;
    struct Rec { int t; hipEvent_t a, b; };
    std::vector<Rec> recs;
    size_t ev_next = 0;
    double t_ms[T_COUNT] = {0};
    long t_n[T_COUNT] = {0};

    hipEvent_t ev() {
        if (ev_next == ev_pool.size()) {
            hipEvent_t e;
            MDP_HIP(hipEventCreate(&e));
            ev_pool.push_back(e);
        }
        return ev_pool[ev_next++];
    }
    template <class F>
    void timed(int t, F&& f) {
        if (!timing || !(timing_mask & (1u << t))) {
            f();
            return;
        }
        hipEvent_t a = ev(), b = ev();
        MDP_HIP(hipEventRecord(a, st));
        f();
        MDP_HIP(hipEventRecord(b, st));
        recs.push_back({t, a, b});
        if (recs.size() > 100000) flush_timers();   // no mid-run synchronisation in practice
    }
    void flush_timers() {
        if (recs.empty()) return;
        MDP_HIP(hipStreamSynchronize(st));
        for (auto& r : recs) {
            float ms = 0;
            MDP_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            t_ms[r.t] += ms;
            t_n[r.t] += 1;
        }
        recs.clear();
        ev_next = 0;
    }
    ~mdp_md() {
        if (st) {
            (void)hipStreamSynchronize(st);
            for (auto e : ev_pool) (void)hipEventDestroy(e);
            (void)hipStreamDestroy(st);
        }
    }

    void init(const mdp_md_params* p, int natoms, const double* box, const double* x0, const double* v0,
              const int* types) {
        if (!p || natoms <= 0 || !box || !x0 || !types) throw Error("mdp_md_create: invalid arguments");
        if (!templated_j(p->twojmax)) throw Error("mdp_md_create: twojmax must be 2, 4, 6 or 8");
        if (p->rebuild_every < 1) throw Error("mdp_md_create: rebuild_every must be >= 1");
        N = natoms;
        ntypes = p->ntypes;
        J = p->twojmax;
        for (int d = 0; d < 3; ++d) {
            L[d] = box[d];
            if (!(L[d] > 0)) throw Error("mdp_md_create: box lengths must be positive");
        }
        for (long k = 0; k < 3L * N; ++k)
            if (!std::isfinite(x0[k]) || (v0 && !std::isfinite(v0[k])))
                throw Error("mdp_md_create: non-finite position or velocity");
        for (int i = 0; i < N; ++i)
            if (types[i] < 1 || types[i] > ntypes) throw Error("mdp_md_create: atom type out of range 1..ntypes");
        dt = p->dt;
        dtf = 0.5 * dt * FTM2V;
        skin = p->skin;
        rebuild_every = p->rebuild_every;
        MDP_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));

        SnapIndex ix;
        ix.build(J, nullptr);
        std::vector<double> radelem(ntypes + 1, 0.0), wjelem(ntypes + 1, 0.0);
        std::vector<int> map(ntypes + 1, 0);
        rcmax = 0;
        for (int t = 1; t <= ntypes; ++t) {
            radelem[t] = p->elemradius[t - 1];
            wjelem[t] = p->elemweight[t - 1];
            rcmax = std::max(rcmax, 2.0 * radelem[t] * p->rcutfac);
        }
        tab.upload(ix, nullptr, nullptr, radelem.data(), wjelem.data(), map.data(), ntypes, st);
        rad_h = radelem;
        wj_h = wjelem;
        tab.view.rcutfac = p->rcutfac;
        tab.view.rfac0 = p->rfac0;
        tab.view.rmin0 = p->rmin0;
        tab.view.switchflag = p->switchflag;
        // chemflag = 0: every atom uses the first coefficient block (cpuSnap.cpp:573, :924)
        const int ncoeffall = ix.idxb_max + 1;
        std::vector<double> bt = y_beta_table(ix, p->coeff, 0, 0, 0, 1, 0, ncoeffall, p->bnormflag);
        e0 = p->coeff[0];
        if (p->bzeroflag) {
            for (int k = 0; k < ix.idxb_max; ++k) {
                const int j = ix.idxb[3 * k + 2];
                const double bz = p->bnormflag ? 1.0 : (j + 1.0);
                e0 -= p->coeff[1 + k] * bz;
            }
        }
        nrows = n_half_rows(J);
        massinv_h.assign(ntypes + 1, 0.0);
        for (int t = 1; t <= ntypes; ++t) massinv_h[t] = 1.0 / p->mass[t - 1];
        rc2_h.assign((ntypes + 1) * (ntypes + 1), 0.0);
        for (int a = 1; a <= ntypes; ++a)
            for (int b = 1; b <= ntypes; ++b) {
                const double rc = (radelem[a] + radelem[b]) * p->rcutfac;
                rc2_h[a * (ntypes + 1) + b] = rc * rc;
            }
        rcnb = rcmax + skin;
        for (int d = 0; d < 3; ++d) ncell[d] = std::max(1, static_cast<int>(std::floor(L[d] / rcnb)));
        use_cells = ncell[0] >= 3 && ncell[1] >= 3 && ncell[2] >= 3;
        if (!use_cells && N > 200000) throw Error("mdp_md_create: box too small for cell lists at this atom count");
        // U and Y of the MD path live in the compact half layout (rows 2mb <= j)
        const size_t n = N, iu = static_cast<size_t>(ix.half_count()),
                     ncells = static_cast<size_t>(ncell[0]) * ncell[1] * ncell[2];
        auto upload = [&](DevBuf& b, const void* h, size_t bytes) {
            b.ensure(bytes);
            MDP_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st));
        };
        upload(x, x0, sizeof(double) * 3 * n);
        std::vector<double> vz(3 * n, 0.0);
        upload(v, v0 ? v0 : vz.data(), sizeof(double) * 3 * n);
        upload(type, types, sizeof(int) * n);
        upload(massinv, massinv_h.data(), sizeof(double) * massinv_h.size());
        std::vector<double> mass_h(ntypes + 1, 0.0);
        for (int t = 1; t <= ntypes; ++t) mass_h[t] = p->mass[t - 1];
        upload(mass_t, mass_h.data(), sizeof(double) * mass_h.size());
        upload(beta_t, bt.data(), sizeof(double) * bt.size());
        upload(rc2, rc2_h.data(), sizeof(double) * rc2_h.size());
        std::vector<int> order(nrows);
        yi_row_order(J, order.data());
        upload(row_order, order.data(), sizeof(int) * order.size());
        f.ensure(sizeof(double) * 3 * n);
        MDP_HIP(hipMemsetAsync(f.p, 0, sizeof(double) * 3 * n, st));
        xref.ensure(sizeof(double) * 3 * n);
        cell_of.ensure(sizeof(int) * n);
        cell_cnt.ensure(sizeof(int) * (ncells + 1));
        cell_start.ensure(sizeof(int) * (ncells + 1));
        cell_fill.ensure(sizeof(int) * (ncells + 1));
        cell_atoms.ensure(sizeof(int) * n);
        scan_tmp.ensure(sizeof(int) * (scan_tmp_ints(static_cast<int>(std::max(n, ncells))) + 8));
        nb_cnt.ensure(sizeof(int) * n);
        nb_j.ensure(sizeof(int) * n * NBMAX);
        nb_s.ensure(sizeof(int) * n * NBMAX);
        in_cnt.ensure(sizeof(int) * n);
        in_e.ensure(sizeof(int) * n * NBMAX);
        pcnt.ensure(sizeof(int) * n);
        poff.ensure(sizeof(int) * (n + 1));
        entry_pair.ensure(sizeof(int) * n * NBMAX);
        Ur.ensure(sizeof(double) * n * iu);
        Ui.ensure(sizeof(double) * n * iu);
        Yr.ensure(sizeof(double) * n * iu);
        Yi.ensure(sizeof(double) * n * iu);
        epart.ensure(sizeof(double) * n * nrows);
        eatom.ensure(sizeof(double) * n);
        vir.ensure(sizeof(double) * 6 * n);
        flags.ensure(sizeof(int) * 4);
        maxd2.ensure(2 * sizeof(unsigned long long));   // [0] max displacement^2 bits, [1] neighbour entries
        pair_sum.ensure(sizeof(unsigned long long));
        MDP_HIP(hipMemsetAsync(pair_sum.p, 0, sizeof(unsigned long long), st));
        MDP_HIP(hipMemsetAsync(Yr.p, 0, sizeof(double) * n * iu, st));
        MDP_HIP(hipMemsetAsync(Yi.p, 0, sizeof(double) * n * iu, st));
        MDP_HIP(hipStreamSynchronize(st));
    }

    void ensure_pairs(size_t cap) {
        if (pi.cap >= sizeof(int) * std::max<size_t>(cap, 1)) return;
        const size_t c = cap + cap / 4 + 64;
        pi.ensure(sizeof(int) * c);
        pj.ensure(sizeof(int) * c);
        pti.ensure(sizeof(int) * c);
        ptj.ensure(sizeof(int) * c);
        prij.ensure(sizeof(double) * 3 * c);
        pgeo.ensure(sizeof(double) * 5 * c);
        dedr.ensure(sizeof(double) * 3 * c);
    }

    // the cell-list neighbour kernel: 0 = eight lanes per atom through L1, 1 = the per-wavefront LDS tile
    void neighbors_launch(int v) {
        const double rcs = rcnb;
        if (v == 1)
            hipLaunchKernelGGL(k_neighbors_tile, dim3(static_cast<unsigned>(ncell[0] * ncell[1] * ncell[2])), dim3(64), 0,
                               st, x.as<double>(), cell_start.as<int>(), cell_atoms.as<int>(), nb_cnt.as<int>(),
                               nb_j.as<int>(), nb_s.as<int>(), L[0], L[1], L[2], ncell[0], ncell[1], ncell[2],
                               rcs * rcs, N, flags.as<int>());
        else
            hipLaunchKernelGGL(k_neighbors_cell, dim3(nb(static_cast<long>(N) * NSUB, 256)), dim3(256), 0, st,
                               x.as<double>(), cell_of.as<int>(), cell_start.as<int>(), cell_atoms.as<int>(),
                               nb_cnt.as<int>(), nb_j.as<int>(), nb_s.as<int>(), L[0], L[1], L[2], ncell[0], ncell[1],
                               ncell[2], rcs * rcs, N, flags.as<int>());
    }

    void rebuild(bool check_disp) {
        const double rcs = rcnb;
        const size_t ncells = static_cast<size_t>(ncell[0]) * ncell[1] * ncell[2];
        MDP_HIP(hipMemsetAsync(cell_cnt.p, 0, sizeof(int) * (ncells + 1), st));
        MDP_HIP(hipMemsetAsync(cell_fill.p, 0, sizeof(int) * (ncells + 1), st));
        MDP_HIP(hipMemsetAsync(flags.p, 0, sizeof(int) * 4, st));
        MDP_HIP(hipMemsetAsync(maxd2.p, 0, 2 * sizeof(unsigned long long), st));
        MDP_HIP(hipMemsetAsync(in_cnt.p, 0, sizeof(int) * N, st));
        hipLaunchKernelGGL(k_wrap_bin, dim3(nb(N, 256)), dim3(256), 0, st, x.as<double>(), xref.as<double>(),
                           cell_of.as<int>(), cell_cnt.as<int>(), maxd2.as<unsigned long long>(), L[0], L[1], L[2],
                           ncell[0], ncell[1], ncell[2], N, check_disp ? 1 : 0, flags.as<int>());
        if (use_cells) {
            exclusive_scan(cell_cnt.as<int>(), cell_start.as<int>(), static_cast<int>(ncells), scan_tmp.as<int>(), st);
            hipLaunchKernelGGL(k_cell_fill, dim3(nb(N, 256)), dim3(256), 0, st, cell_of.as<int>(), cell_start.as<int>(),
                               cell_fill.as<int>(), cell_atoms.as<int>(), N);
            hipLaunchKernelGGL(k_cell_sort, dim3(nb(static_cast<long>(ncells), 256)), dim3(256), 0, st,
                               cell_start.as<int>(), cell_atoms.as<int>(), static_cast<int>(ncells), flags.as<int>());
            neighbors_launch(nb_variant);
        } else {
            hipLaunchKernelGGL(k_neighbors_brute, dim3(nb(N, 64)), dim3(64), 0, st, x.as<double>(), nb_cnt.as<int>(),
                               nb_j.as<int>(), nb_s.as<int>(), L[0], L[1], L[2], rcs, rcs * rcs, N, flags.as<int>());
        }
        hipLaunchKernelGGL(k_incoming, dim3(nb(N, 256)), dim3(256), 0, st, nb_cnt.as<int>(), nb_j.as<int>(),
                           in_cnt.as<int>(), in_e.as<int>(), N, flags.as<int>());
        hipLaunchKernelGGL(k_incoming_sort, dim3(nb(N, 256)), dim3(256), 0, st, in_cnt.as<int>(), in_e.as<int>(), N);
        hipLaunchKernelGGL(k_sum_counts, dim3(nb(N, 256)), dim3(256), 0, st, nb_cnt.as<int>(), N,
                           maxd2.as<unsigned long long>() + 1);
        // a single small readback per rebuild: overflow flags, displacement and entry count
        int fl[4];
        unsigned long long md2[2] = {0, 0};
        MDP_HIP(hipMemcpyAsync(fl, flags.p, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
        MDP_HIP(hipMemcpyAsync(md2, maxd2.p, sizeof(md2), hipMemcpyDeviceToHost, st));
        MDP_HIP(hipStreamSynchronize(st));
        if (fl[0]) {
            stats[5] += 1;
            throw Error(std::string("neighbour build failed:") + ((fl[0] & 1) ? " >96 neighbours per atom" : "") +
                        ((fl[0] & 2) ? " >48 atoms per cell" : "") + ((fl[0] & 4) ? " >96 incoming entries" : "") +
                        ((fl[0] & 8) ? " non-finite or runaway positions" : ""));
        }
        double d2;
        std::memcpy(&d2, &md2[0], sizeof(d2));
        stats[2] += 1;
        if (check_disp && d2 > 0.25 * skin * skin) stats[3] += 1;
        // entry count bounds the pair count until the next rebuild
        const long tot = static_cast<long>(md2[1]);
        stats[1] = tot;
        ensure_pairs(static_cast<size_t>(tot));
    }

    void pairs() {
        hipLaunchKernelGGL(k_pair_count_md, dim3(nb(static_cast<long>(N) * PSUB, 256)), dim3(256), 0, st, x.as<double>(), type.as<int>(),
                           nb_cnt.as<int>(), nb_j.as<int>(), nb_s.as<int>(), rc2.as<double>(), ntypes, pcnt.as<int>(),
                           L[0], L[1], L[2], N);
        exclusive_scan(pcnt.as<int>(), poff.as<int>(), N, scan_tmp.as<int>(), st);
        hipLaunchKernelGGL(k_pair_fill_md, dim3(nb(static_cast<long>(N) * PSUB, 256)), dim3(256), 0, st, x.as<double>(), type.as<int>(),
                           nb_cnt.as<int>(), nb_j.as<int>(), nb_s.as<int>(), rc2.as<double>(), ntypes, poff.as<int>(),
                           pi.as<int>(), pj.as<int>(), pti.as<int>(), ptj.as<int>(), prij.as<double>(),
                           entry_pair.as<int>(), L[0], L[1], L[2], N, tab.view, pgeo.as<double>());
    }

    void dedr_launch(int variant) {
        const SnapDev& s = tab.view;
        if (variant == 5) {
            yt.ensure(sizeof(double) * 2 * static_cast<size_t>(tab.idx.half_count()) * N);
            launch_dedr_atom(s, Yr.as<double>(), Yi.as<double>(), yt.as<double>(), prij.as<double>(), pti.as<int>(),
                             ptj.as<int>(), poff.as<int>(), dedr.as<double>(), N, st);
        } else {
            launch_dedr(s, Yr.as<double>(), Yi.as<double>(), prij.as<double>(), pi.as<int>(), pti.as<int>(),
                        ptj.as<int>(), nullptr, dedr.as<double>(), N, static_cast<int>(stats[1]), poff.as<int>() + N,
                        st, true, variant);
        }
    }

    // reference pair potential of every atom into zacc, added by the tally that follows
    void pair_potential() {
        if (!zbl) return;
        timed(T_PAIRPOT, [&] {
            hipLaunchKernelGGL(k_zbl_md, dim3(nb(static_cast<long>(N) * PSUB, 256)), dim3(256), 0, st, x.as<double>(),
                               nb_cnt.as<int>(), nb_j.as<int>(), nb_s.as<int>(), L[0], L[1], L[2], N, zp,
                               zacc.as<double>());
        });
    }

    // Adds the ZBL pair potential (mu = [cut_inner, cut_outer, z]) to the forces from the next force
    // evaluation on.  The neighbour list is rebuilt at max(SNAP cutoff, cut_outer) + skin, so that the
    // pair potential's pairs stay in the list between rebuilds exactly as SNAP's do.
    void set_zbl(double cut_inner, double cut_outer, double z) {
        if (!(cut_outer > cut_inner) || !(cut_inner > 0.0) || !(z > 0.0))
            throw Error("mdp_md_set_zbl: need 0 < cut_inner < cut_outer and z > 0");
        const double rcs = std::max(rcmax, cut_outer) + skin;
        int nc[3];
        for (int d = 0; d < 3; ++d) nc[d] = std::max(1, static_cast<int>(std::floor(L[d] / rcs)));
        const bool cells = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3;
        if (!cells && N > 200000) throw Error("mdp_md_set_zbl: box too small for cell lists at this atom count");
        const size_t ncells = static_cast<size_t>(nc[0]) * nc[1] * nc[2];
        cell_cnt.ensure(sizeof(int) * (ncells + 1));
        cell_start.ensure(sizeof(int) * (ncells + 1));
        cell_fill.ensure(sizeof(int) * (ncells + 1));
        scan_tmp.ensure(sizeof(int) * (scan_tmp_ints(static_cast<int>(std::max(static_cast<size_t>(N), ncells))) + 8));
        zacc.ensure(sizeof(double) * 10 * static_cast<size_t>(N));
        for (int d = 0; d < 3; ++d) ncell[d] = nc[d];
        use_cells = cells;
        rcnb = rcs;
        zp = zbl_params(cut_inner, cut_outer, z);
        zbl_rout = cut_outer;
        zbl = true;
        forces_current = false;
    }

    void ui_launch(int v) {
        if (v == 0)
            launch_ui_geo(tab.view, Ur.as<double>(), Ui.as<double>(), pgeo.as<double>(), poff.as<int>(), N, 1.0, st);
        else
            launch_ui_geo_lds(tab.view, Ur.as<double>(), Ui.as<double>(), pgeo.as<double>(), poff.as<int>(), N, 1.0,
                              v, st);
    }

    // plain SNAP forces; outputs to (fo, vo, eo), by default the engine's f, vir and eatom
    void snap_and_tally(bool integrate, double* fo = nullptr, double* vo = nullptr, double* eo = nullptr) {
        const SnapDev& s = tab.view;
        const bool count = (fo == nullptr);   // audit evaluations do not count as steps
        if (count) ++force_evals;
        if (!fo) fo = f.as<double>();
        if (!vo) vo = vir.as<double>();
        if (!eo) eo = eatom.as<double>();
        timed(T_UI, [&] { ui_launch(ui_variant); });
        timed(T_YI, [&] {
            launch_yi_half(s, Ur.as<double>(), Ui.as<double>(), beta_t.as<double>(), Yr.as<double>(), Yi.as<double>(),
                           epart.as<double>(), row_order.as<int>(), N, st, yi_waves);
        });
        timed(T_DEDR, [&] {
            dedr_launch(dedr_variant);
        });
        pair_potential();
        timed(T_TALLY, [&] {
            hipLaunchKernelGGL(k_tally, dim3(nb(static_cast<long>(N) * PSUB, 256)), dim3(256), 0, st, dedr.as<double>(), prij.as<double>(),
                               poff.as<int>(), in_cnt.as<int>(), in_e.as<int>(), entry_pair.as<int>(),
                               epart.as<double>(), nrows, e0, nullptr, fo, vo, eo, v.as<double>(), type.as<int>(),
                               massinv.as<double>(), dtf, integrate ? 1 : 0, N, count ? pair_sum.as<unsigned long long>() : nullptr,
                               zbl ? zacc.as<double>() : nullptr);
        });
    }

    void set_surrogate(const float* params, const double* aff, int nin, int js, int every, double tol) {
        if (!params || !aff) throw Error("mdp_md_set_surrogate: null parameters");
        if (js != 2 && js != 4) throw Error("mdp_md_set_surrogate: descriptor twojmax must be 2 or 4");
        if (every < 0) throw Error("mdp_md_set_surrogate: audit_every must be >= 0");
        SnapIndex ix;
        ix.build(js, nullptr);
        if (nin != ix.idxb_max)
            throw Error("mdp_md_set_surrogate: nin must be the number of bispectrum components of twojmax");
        std::vector<int> map(ntypes + 1, 0);
        tab_s.upload(ix, nullptr, nullptr, rad_h.data(), wj_h.data(), map.data(), ntypes, st);
        tab_s.view.rcutfac = tab.view.rcutfac;
        tab_s.view.rfac0 = tab.view.rfac0;
        tab_s.view.rmin0 = tab.view.rmin0;
        tab_s.view.switchflag = tab.view.switchflag;
        // Y weights of each coupling triple as (bispectrum index, multiplicity factor)
        std::vector<int> jjb;
        std::vector<double> fac;
        y_triple_weights(ix, 0, jjb, fac);
        sur_ntrip = static_cast<int>(jjb.size());
        // the model's descriptors are B - bzero_j (wself = 1): fold bzero into the centring
        std::vector<double> a(aff, aff + 34);
        for (int k = 0; k < nin; ++k) a[k] += ix.idxb[3 * k + 2] + 1.0;
        auto put = [&](DevBuf& b, const void* h, size_t bytes) {
            b.ensure(bytes);
            MDP_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, st));
        };
        put(sParams, params, sizeof(float) * MLP_NPARAM);
        put(sAff, a.data(), sizeof(double) * 34);
        put(sJjb, jjb.data(), sizeof(int) * jjb.size());
        std::vector<int> sorder((js / 2 + 1) * (js / 2 + 1));   // half-U rows (j, mb), 2mb <= j <= js
        yi_row_order(js, sorder.data());
        put(sRowOrder, sorder.data(), sizeof(int) * sorder.size());
        put(sFac, fac.data(), sizeof(double) * fac.size());
        sPacked.ensure(sizeof(unsigned short) * MLP_NPACK);
        launch_mlp_pack(sParams.as<float>(), sPacked.as<unsigned short>(), st);
        const size_t n = N, iu = ix.idxu_max;
        sUr.ensure(sizeof(double) * n * iu);
        sUi.ensure(sizeof(double) * n * iu);
        sYr.ensure(sizeof(double) * n * iu);
        sYi.ensure(sizeof(double) * n * iu);
        sB.ensure(sizeof(double) * n * nin);
        sG.ensure(sizeof(double) * n * nin);
        sE.ensure(sizeof(double) * n);
        sBeta.ensure(sizeof(double) * n * sur_ntrip);
        fP.ensure(sizeof(double) * 3 * n);
        vP.ensure(sizeof(double) * 6 * n);
        eP.ensure(sizeof(double) * n);
        ssq.ensure(sizeof(double) * 256);
        MDP_HIP(hipStreamSynchronize(st));
        sur_nin = nin;
        audit_every = every;
        audit_tol = tol;
        audits = fallbacks = 0;
        rmse_last = rmse_max = rmse_sum = 0;
        sur = true;
        forces_current = false;
    }

    // surrogate forces: Utot and B at the descriptor twojmax, the MLP's energy and dE/dB per atom, then
    // dE/dr by the adjoint route (Y with per-atom weights dE/dB, then the dU/dr contraction)
    void sur_and_tally(bool integrate) {
        const SnapDev& s4 = tab_s.view;
        ++force_evals;
        timed(T_SUR, [&] {
            launch_ui_geo(s4, sUr.as<double>(), sUi.as<double>(), pgeo.as<double>(), poff.as<int>(), N, 1.0, st);
            launch_b_half(s4, sUr.as<double>(), sUi.as<double>(), sB.as<double>(), N, 0, st);
            launch_mlp_eval(sPacked.as<unsigned short>(), sParams.as<float>(), sAff.as<double>(), sur_nin,
                            sB.as<double>(), N, sE.as<double>(), sG.as<double>(), st);
            launch_beta_peratom(sG.as<double>(), sJjb.as<int>(), sFac.as<double>(), sur_ntrip, N, sBeta.as<double>(),
                                st);
            launch_yi_half_pa(s4, sUr.as<double>(), sUi.as<double>(), sBeta.as<double>(), sYr.as<double>(),
                              sYi.as<double>(), sRowOrder.as<int>(), N, st);
            launch_dedr(s4, sYr.as<double>(), sYi.as<double>(), prij.as<double>(), pi.as<int>(), pti.as<int>(),
                        ptj.as<int>(), nullptr, dedr.as<double>(), N, static_cast<int>(stats[1]), poff.as<int>() + N, st,
                        true, 10);
        });
        pair_potential();
        timed(T_TALLY, [&] {
            hipLaunchKernelGGL(k_tally, dim3(nb(static_cast<long>(N) * PSUB, 256)), dim3(256), 0, st, dedr.as<double>(), prij.as<double>(),
                               poff.as<int>(), in_cnt.as<int>(), in_e.as<int>(), entry_pair.as<int>(),
                               epart.as<double>(), 0, 0.0, sE.as<double>(), f.as<double>(), vir.as<double>(),
                               eatom.as<double>(), v.as<double>(), type.as<int>(), massinv.as<double>(), dtf,
                               integrate ? 1 : 0, N, pair_sum.as<unsigned long long>(), zbl ? zacc.as<double>() : nullptr);
        });
    }

    // plain SNAP on the same state; RMS force difference; plain forces replace the surrogate's when the
    // difference exceeds audit_tol; then the second velocity half-step
    void audit() {
        snap_and_tally(false, fP.as<double>(), vP.as<double>(), eP.as<double>());
        timed(T_AUDIT, [&] {
            hipLaunchKernelGGL(k_diff_sumsq, dim3(256), dim3(256), 0, st, f.as<double>(), fP.as<double>(),
                               3L * N, ssq.as<double>());
        });
        double part[256];
        MDP_HIP(hipMemcpyAsync(part, ssq.p, sizeof(part), hipMemcpyDeviceToHost, st));
        MDP_HIP(hipStreamSynchronize(st));
        double sum = 0.0;
        for (double v_ : part) sum += v_;
        const double rmse = std::sqrt(sum / (3.0 * N));
        audits += 1;
        rmse_last = rmse;
        rmse_max = std::max(rmse_max, rmse);
        rmse_sum += rmse;
        if (!(rmse <= audit_tol)) {
            fallbacks += 1;
            MDP_HIP(hipMemcpyAsync(f.p, fP.p, sizeof(double) * 3 * N, hipMemcpyDeviceToDevice, st));
            MDP_HIP(hipMemcpyAsync(vir.p, vP.p, sizeof(double) * 6 * N, hipMemcpyDeviceToDevice, st));
            MDP_HIP(hipMemcpyAsync(eatom.p, eP.p, sizeof(double) * N, hipMemcpyDeviceToDevice, st));
        }
        hipLaunchKernelGGL(k_vv2, dim3(nb(N, 256)), dim3(256), 0, st, v.as<double>(), f.as<double>(), type.as<int>(),
                           massinv.as<double>(), dtf, N);
    }

    // Time every compiled variant of the two heaviest kernels on the live data (both are pure
    // functions of their inputs, so re-running them is harmless) and keep the fastest.
    template <class F>
    double time_launch(F&& f, int reps) {
        hipEvent_t a = ev(), b = ev();
        f();   // warm-up (code object load, caches)
        MDP_HIP(hipEventRecord(a, st));
        for (int r = 0; r < reps; ++r) f();
        MDP_HIP(hipEventRecord(b, st));
        MDP_HIP(hipEventSynchronize(b));
        float ms = 0;
        MDP_HIP(hipEventElapsedTime(&ms, a, b));
        ev_next -= 2;
        return ms / reps;
    }
    void autotune() {
        const SnapDev& s = tab.view;
        const char* env = std::getenv("MDP_AUTOTUNE");
        if (env && std::string(env) == "0") {
            tuned = true;
            return;
        }
        std::string js = "{";
        int ids[16];
        int nv = dedr_variants(ids);
        // MDP_DEDR_VARIANT pins the dE/dr kernel variant (one of dedr_variants) instead of timing them all
        const char* pin = std::getenv("MDP_DEDR_VARIANT");
        if (pin) {
            const int v = std::atoi(pin);
            bool known = false;
            for (int k = 0; k < nv; ++k) known |= (ids[k] == v);
            if (!known) throw Error("MDP_DEDR_VARIANT is not a dE/dr kernel variant");
            ids[0] = v;
            nv = 1;
        }
        double best = 1e30, best_xcd = 1e30;
        int pick_xcd = -1;
        js += "\"compute_dedr\": {";
        for (int k = 0; k < nv; ++k) {
            const double t = time_launch([&] {
                dedr_launch(ids[k]);
            }, 3);
            js += (k ? ", " : "") + std::string("\"") + std::to_string(ids[k]) + "\": " + std::to_string(t);
            if (t < best) { best = t; dedr_variant = ids[k]; }
            if ((ids[k] == 12 || ids[k] == 13) && t < best_xcd) { best_xcd = t; pick_xcd = ids[k]; }
        }
        // Variants 12 and 13 give each XCD one contiguous range of pairs, so the Y rows of an atom are filled
        // into one L2 only: they move about a third of the HBM bytes of the other mappings.  Within 3% of the
        // fastest time they are preferred.
        if (pick_xcd >= 0 && best_xcd <= 1.03 * best) dedr_variant = pick_xcd;
        js += "}, \"compute_yi\": {";
        nv = yi_half_variants(J, ids);
        // MDP_YI_VARIANT pins compute_yi to one of yi_half_variants
        if (const char* yp = std::getenv("MDP_YI_VARIANT")) {
            const int v = std::atoi(yp);
            bool known = false;
            for (int k = 0; k < nv; ++k) known |= (ids[k] == v);
            if (!known) throw Error("MDP_YI_VARIANT is not a compute_yi variant");
            ids[0] = v;
            nv = 1;
        }
        best = 1e30;
        for (int k = 0; k < nv; ++k) {
            const double t = time_launch([&] {
                launch_yi_half(s, Ur.as<double>(), Ui.as<double>(), beta_t.as<double>(), Yr.as<double>(),
                               Yi.as<double>(), epart.as<double>(), row_order.as<int>(), N, st, ids[k]);
            }, 3);
            js += (k ? ", " : "") + std::string("\"") + std::to_string(ids[k]) + "\": " + std::to_string(t);
            if (t < best) { best = t; yi_waves = ids[k]; }
        }
        // compute_ui: the per-atom kernel, or the LDS-sum kernel on 256, 512 or 1024 workgroups;
        // MDP_UI_VARIANT pins one of 0, 256, 512, 1024
        js += "}, \"compute_ui\": {";
        int uids[4] = {0, 256, 512, 1024}, nu = 4;
        if (const char* up = std::getenv("MDP_UI_VARIANT")) {
            const int v = std::atoi(up);
            if (v != 0 && v != 256 && v != 512 && v != 1024)
                throw Error("MDP_UI_VARIANT is not a compute_ui variant (0, 256, 512 or 1024)");
            uids[0] = v;
            nu = 1;
        }
        best = 1e30;
        for (int k = 0; k < nu; ++k) {
            const double t = time_launch([&] { ui_launch(uids[k]); }, 3);
            js += (k ? ", " : "") + std::string("\"") + std::to_string(uids[k]) + "\": " + std::to_string(t);
            if (t < best) { best = t; ui_variant = uids[k]; }
        }
        // cell-list neighbour kernel (when the box has cell lists): both write the same lists from the
        // cell arrays of the last rebuild; MDP_NB_VARIANT pins 0 or 1
        if (use_cells) {
            js += "}, \"neighbor_build\": {";
            int nids[2] = {0, 1}, nn = 2;
            if (const char* np_ = std::getenv("MDP_NB_VARIANT")) {
                const int v = std::atoi(np_);
                if (v != 0 && v != 1) throw Error("MDP_NB_VARIANT is not a neighbour kernel variant (0 or 1)");
                nids[0] = v;
                nn = 1;
            }
            best = 1e30;
            for (int k = 0; k < nn; ++k) {
                const double t = time_launch([&] { neighbors_launch(nids[k]); }, 3);
                js += (k ? ", " : "") + std::string("\"") + std::to_string(nids[k]) + "\": " + std::to_string(t);
                if (t < best) { best = t; nb_variant = nids[k]; }
            }
            neighbors_launch(nb_variant);
        }
        js += "}, \"chosen\": {\"compute_dedr\": " + std::to_string(dedr_variant) +
              ", \"compute_yi_waves\": " + std::to_string(yi_waves) + ", \"compute_ui\": " +
              std::to_string(ui_variant) + ", \"neighbor_build\": " + std::to_string(nb_variant) + "}}";
        tuning_json = js;
        tuned = true;
    }

    void compute() {
        timed(T_REBUILD, [&] { rebuild(false); });
        timed(T_PAIRS, [&] { pairs(); });
        snap_and_tally(false);
        if (!tuned) {
            autotune();
            snap_and_tally(false);   // final state from the chosen variants
        }
        if (sur) sur_and_tally(false);
        MDP_HIP(hipGetLastError());
        MDP_HIP(hipStreamSynchronize(st));
        int np = 0;
        MDP_HIP(hipMemcpy(&np, poff.as<int>() + N, sizeof(int), hipMemcpyDeviceToHost));
        stats[0] = np;
        forces_current = true;
    }

    // ThermoOutput of the upstream app (Calculation.cpp:436-465) for the current state: potential energy
    // per atom, temperature (tdof = 3 (N - 1)), kinetic energy per atom, pressure (bar, metal units).
    void thermo(double* out) {
        if (!forces_current) compute();
        const int nparts = nb(N, 256);
        kepart.ensure(sizeof(double) * 5 * nparts);
        hipLaunchKernelGGL(k_thermo_part, dim3(nparts), dim3(256), 0, st, eatom.as<double>(), vir.as<double>(),
                           v.as<double>(), type.as<int>(), mass_t.as<double>(), N, kepart.as<double>());
        std::vector<double> h(5 * static_cast<size_t>(nparts));
        MDP_HIP(hipMemcpyAsync(h.data(), kepart.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        MDP_HIP(hipStreamSynchronize(st));
        double q[5] = {0, 0, 0, 0, 0};
        for (int b = 0; b < nparts; ++b)
            for (int k = 0; k < 5; ++k) q[k] += h[5 * static_cast<size_t>(b) + k];
        constexpr double NKTV2P = 1.6021765e6;
        const double tdof = 3.0 * (N - 1), tfactor = MVV2E / (tdof * BOLTZ);
        const double temp = tfactor * q[4];
        out[0] = q[0] / N;
        out[1] = temp;
        out[2] = 0.5 * tdof * BOLTZ * temp / N;
        out[3] = (tdof * BOLTZ * temp + q[1] + q[2] + q[3]) / 3.0 * (1.0 / (L[0] * L[1] * L[2])) * NKTV2P;
    }

    // One thermostat half-step on the device: block sums of m v^2, then the chain (one thread).
    void nh_half(double t_target) {
        const int nparts = nb(N, 256);
        hipLaunchKernelGGL(k_ke_part, dim3(nparts), dim3(256), 0, st, v.as<double>(), type.as<int>(), mass_t.as<double>(),
                           N, kepart.as<double>());
        const double tfactor = MVV2E / (nvt_tdof * BOLTZ);
        hipLaunchKernelGGL(k_nh_chain, dim3(1), dim3(64), 0, st, kepart.as<double>(), nparts, nh.as<double>(), t_target,
                           nvt_tfreq, nvt_tdof, BOLTZ, nvt_drag, dt, tfactor, nvt_mtchain, nvt_nc, nvt_massflag);
    }

    // new positions (and velocities, when given) from the host; the next step rebuilds and recomputes
    void set_state(const double* xh, const double* vh) {
        if (!xh) throw Error("mdp_md_set_state: null positions");
        for (long k = 0; k < 3L * N; ++k)
            if (!std::isfinite(xh[k]) || (vh && !std::isfinite(vh[k])))
                throw Error("mdp_md_set_state: non-finite position or velocity");
        MDP_HIP(hipMemcpyAsync(x.p, xh, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st));
        if (vh) MDP_HIP(hipMemcpyAsync(v.p, vh, sizeof(double) * 3 * N, hipMemcpyHostToDevice, st));
        MDP_HIP(hipStreamSynchronize(st));
        forces_current = false;
    }

    void set_nvt(double t_start, double t_stop, double t_period, double drag, double tdof, int mtchain, int nc_tchain,
                 long beginstep, long endstep, long currentstep) {
        if (!(t_period > 0.0)) throw Error("mdp_md_set_nvt: the thermostat period must be positive");
        if (mtchain < 1 || mtchain > NH_MAX - 1) throw Error("mdp_md_set_nvt: mtchain must be 1..9");
        if (nc_tchain < 1) throw Error("mdp_md_set_nvt: nc_tchain must be positive");
        if (endstep <= beginstep) throw Error("mdp_md_set_nvt: endstep must be above beginstep");
        if (!(t_start > 0.0) || !(t_stop > 0.0)) throw Error("mdp_md_set_nvt: temperatures must be positive");
        nvt = true;
        nvt_tstart = t_start;
        nvt_tstop = t_stop;
        nvt_tfreq = 1.0 / t_period;   // readdata.cpp:432
        nvt_drag = drag;
        nvt_tdof = tdof > 0.0 ? tdof : 3.0 * (N - 1);   // Calculation.cpp:450
        nvt_mtchain = mtchain;
        nvt_nc = nc_tchain;
        nvt_begin = beginstep;
        nvt_end = endstep;
        nvt_step = currentstep;
        nh.ensure(sizeof(double) * NH_SIZE);
        kepart.ensure(sizeof(double) * nb(N, 256));
        std::vector<double> z(NH_SIZE, 0.0);
        z[NH_FACTOR] = 1.0;
        MDP_HIP(hipMemcpyAsync(nh.p, z.data(), sizeof(double) * NH_SIZE, hipMemcpyHostToDevice, st));
        MDP_HIP(hipStreamSynchronize(st));
    }

    void run(int nsteps) {
        if (!forces_current) compute();
        for (int s = 0; s < nsteps; ++s) {
            double t_target = 0.0;
            if (nvt) {   // target ramp of cpuNVTInitialIntegrate (cpuVelocityIntegration.cpp:611-612)
                ++nvt_step;
                const double delta = static_cast<double>(nvt_step - nvt_begin) / static_cast<double>(nvt_end - nvt_begin);
                t_target = nvt_tstart + delta * (nvt_tstop - nvt_tstart);
            }
            timed(T_VV1, [&] {
                if (nvt) {
                    nh_half(t_target);
                    hipLaunchKernelGGL(k_vv1_nh, dim3(nb(N, 256)), dim3(256), 0, st, x.as<double>(), v.as<double>(),
                                       f.as<double>(), type.as<int>(), massinv.as<double>(), nh.as<double>(), dtf, dt,
                                       N);
                } else {
                    hipLaunchKernelGGL(k_vv1, dim3(nb(N, 256)), dim3(256), 0, st, x.as<double>(), v.as<double>(),
                                       f.as<double>(), type.as<int>(), massinv.as<double>(), dtf, dt, N);
                }
            });
            stats[4] += 1;
            if (stats[4] % rebuild_every == 0) timed(T_REBUILD, [&] { rebuild(true); });
            timed(T_PAIRS, [&] { pairs(); });
            if (sur) {
                const bool aud = audit_every > 0 && stats[4] % audit_every == 0;
                sur_and_tally(!aud);
                if (aud) audit();
            } else {
                snap_and_tally(true);
            }
            if (nvt)   // second thermostat half-step after the final velocity update (cpuNVTFinalIntegrate)
                timed(T_VV1, [&] {
                    nh_half(t_target);
                    hipLaunchKernelGGL(k_vscale_nh, dim3(nb(3L * N, 256)), dim3(256), 0, st, v.as<double>(),
                                       nh.as<double>(), N);
                });
        }
        MDP_HIP(hipGetLastError());
        MDP_HIP(hipStreamSynchronize(st));
        int np = 0;
        MDP_HIP(hipMemcpy(&np, poff.as<int>() + N, sizeof(int), hipMemcpyDeviceToHost));
        stats[0] = np;
    }
};

namespace {
template <class F>
int api(F&& f) {
    clear_error();
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 1;
    }
}
}  // namespace

extern "C" {

int mdp_md_set_state(mdp_md* md, const double* x, const double* v) {
    return api([&] {
        if (!md) throw Error("mdp_md_set_state: null handle");
        md->set_state(x, v);
    });
}

int mdp_md_set_zbl(mdp_md* md, double cut_inner, double cut_outer, double z) {
    return api([&] {
        if (!md) throw Error("mdp_md_set_zbl: null handle");
        md->set_zbl(cut_inner, cut_outer, z);
    });
}

int mdp_md_set_nvt(mdp_md* md, double t_start, double t_stop, double t_period, double drag, double tdof, int mtchain,
                   int nc_tchain, long beginstep, long endstep, long currentstep) {
    return api([&] {
        if (!md) throw Error("mdp_md_set_nvt: null handle");
        md->set_nvt(t_start, t_stop, t_period, drag, tdof, mtchain, nc_tchain, beginstep, endstep, currentstep);
    });
}

int mdp_md_thermo(mdp_md* md, double* out4) {
    return api([&] {
        if (!md || !out4) throw Error("mdp_md_thermo: null argument");
        md->thermo(out4);
    });
}

int mdp_md_clear_nvt(mdp_md* md) {
    return api([&] {
        if (!md) throw Error("mdp_md_clear_nvt: null handle");
        md->nvt = false;
    });
}

int mdp_md_nvt_state(mdp_md* md, double* out) {
    return api([&] {
        if (!md || !out) throw Error("mdp_md_nvt_state: null argument");
        if (!md->nvt) throw Error("mdp_md_nvt_state: NVT is not set");
        std::vector<double> h(NH_SIZE);
        MDP_HIP(hipStreamSynchronize(md->st));
        MDP_HIP(hipMemcpy(h.data(), md->nh.p, sizeof(double) * NH_SIZE, hipMemcpyDeviceToHost));
        for (int k = 0; k < NH_MAX; ++k) {
            out[k] = h[NH_MASS + k];
            out[NH_MAX + k] = h[NH_ETA + k];
            out[2 * NH_MAX + k] = h[NH_DOT + k];
            out[3 * NH_MAX + k] = h[NH_DDOT + k];
        }
        out[4 * NH_MAX] = h[NH_T];
        out[4 * NH_MAX + 1] = h[NH_TTARGET];
        out[4 * NH_MAX + 2] = static_cast<double>(md->nvt_step);
    });
}

int mdp_md_create(mdp_md** md, const mdp_md_params* p, int natoms, const double* box, const double* x, const double* v,
                  const int* types) {
    return api([&] {
        if (!md) throw Error("mdp_md_create: null handle pointer");
        *md = nullptr;
        auto h = std::make_unique<mdp_md>();
        h->init(p, natoms, box, x, v, types);
        *md = h.release();
    });
}

int mdp_md_destroy(mdp_md* md) {
    return api([&] { delete md; });
}

int mdp_md_compute(mdp_md* md) {
    return api([&] { md->compute(); });
}

int mdp_md_run(mdp_md* md, int nsteps) {
    return api([&] { md->run(nsteps); });
}

int mdp_md_get(mdp_md* md, double* x, double* v, double* f, double* eatom, double* vatom) {
    return api([&] {
        const size_t n = md->N;
        MDP_HIP(hipStreamSynchronize(md->st));
        if (x) MDP_HIP(hipMemcpy(x, md->x.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        if (v) MDP_HIP(hipMemcpy(v, md->v.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        if (f) MDP_HIP(hipMemcpy(f, md->f.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        if (eatom) MDP_HIP(hipMemcpy(eatom, md->eatom.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        if (vatom) MDP_HIP(hipMemcpy(vatom, md->vir.p, sizeof(double) * 6 * n, hipMemcpyDeviceToHost));
    });
}

int mdp_md_neighbors(mdp_md* md, int* off, int* nbj, int* nbshift, long* nentries) {
    return api([&] {
        const int N = md->N;
        MDP_HIP(hipStreamSynchronize(md->st));
        std::vector<int> cnt(N), js(static_cast<size_t>(N) * NBMAX), ss(static_cast<size_t>(N) * NBMAX);
        MDP_HIP(hipMemcpy(cnt.data(), md->nb_cnt.p, sizeof(int) * N, hipMemcpyDeviceToHost));
        long tot = 0;
        for (int c : cnt) tot += c;
        if (nentries) *nentries = tot;
        if (!nbj) return;
        MDP_HIP(hipMemcpy(js.data(), md->nb_j.p, sizeof(int) * js.size(), hipMemcpyDeviceToHost));
        MDP_HIP(hipMemcpy(ss.data(), md->nb_s.p, sizeof(int) * ss.size(), hipMemcpyDeviceToHost));
        long k = 0;
        for (int i = 0; i < N; ++i) {
            off[i] = static_cast<int>(k);
            for (int m = 0; m < cnt[i]; ++m, ++k) {
                nbj[k] = js[static_cast<size_t>(m) * N + i];
                int sx, sy, sz;
                shift_decode(ss[static_cast<size_t>(m) * N + i], sx, sy, sz);
                nbshift[3 * k] = sx;
                nbshift[3 * k + 1] = sy;
                nbshift[3 * k + 2] = sz;
            }
        }
        off[N] = static_cast<int>(k);
    });
}

int mdp_md_stats(mdp_md* md, long* out6) {
    return api([&] {
        for (int k = 0; k < 6; ++k) out6[k] = md->stats[k];
    });
}

int mdp_md_set_timing(mdp_md* md, int on) {
    return api([&] {
        md->flush_timers();
        MDP_HIP(hipMemsetAsync(md->pair_sum.p, 0, sizeof(unsigned long long), md->st));
        md->pair_steps0 = md->force_evals;
        if (on < 0 || on > 2) throw Error("mdp_md_set_timing: on must be 0, 1 or 2");
        md->timing = on != 0;
        // 2: only the SNAP compute kernels (fewer events inside a timed run)
        md->timing_mask = on == 2 ? (1u << T_UI) | (1u << T_YI) | (1u << T_DEDR) : ~0u;
        for (int k = 0; k < T_COUNT; ++k) {
            md->t_ms[k] = 0;
            md->t_n[k] = 0;
        }
    });
}

int mdp_md_timers(mdp_md* md, double* ms_total, long* launches, int n) {
    return api([&] {
        md->flush_timers();
        for (int k = 0; k < n && k < T_COUNT; ++k) {
            ms_total[k] = md->t_ms[k];
            launches[k] = md->t_n[k];
        }
    });
}

int mdp_md_tuning(mdp_md* md, char* buf, int len) {
    return api([&] {
        if (!buf || len <= 0) return;
        const std::string& j = md->tuning_json;
        const size_t n = std::min(static_cast<size_t>(len - 1), j.size());
        std::memcpy(buf, j.data(), n);
        buf[n] = '\0';
    });
}

int mdp_md_set_surrogate(mdp_md* md, const float* params, const double* aff, int nin, int twojmax, int audit_every,
                         double audit_tol) {
    return api([&] { md->set_surrogate(params, aff, nin, twojmax, audit_every, audit_tol); });
}

int mdp_md_clear_surrogate(mdp_md* md) {
    return api([&] {
        md->sur = false;
        md->forces_current = false;
    });
}

int mdp_md_mean_pairs(mdp_md* md, double* mean) {
    return api([&] {
        unsigned long long sum = 0;
        MDP_HIP(hipMemcpyAsync(&sum, md->pair_sum.p, sizeof(sum), hipMemcpyDeviceToHost, md->st));
        MDP_HIP(hipStreamSynchronize(md->st));
        const long n = md->force_evals - md->pair_steps0;
        *mean = n > 0 ? static_cast<double>(sum) / n : 0.0;
    });
}

int mdp_md_audits(mdp_md* md, double* out5) {
    return api([&] {
        out5[0] = static_cast<double>(md->audits);
        out5[1] = static_cast<double>(md->fallbacks);
        out5[2] = md->rmse_last;
        out5[3] = md->rmse_max;
        out5[4] = md->audits ? md->rmse_sum / md->audits : 0.0;
    });
}

const char* mdp_md_timer_name(int k) { return (k >= 0 && k < T_COUNT) ? kTimerNames[k] : ""; }
int mdp_md_num_timers(void) { return T_COUNT; }

int mdp_md_energy(mdp_md* md, double* epot, double* ekin) {
    return api([&] {
        const size_t n = md->N;
        std::vector<double> e(n), v(3 * n);
        std::vector<int> t(n);
        MDP_HIP(hipStreamSynchronize(md->st));
        MDP_HIP(hipMemcpy(e.data(), md->eatom.p, sizeof(double) * n, hipMemcpyDeviceToHost));
        MDP_HIP(hipMemcpy(v.data(), md->v.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
        MDP_HIP(hipMemcpy(t.data(), md->type.p, sizeof(int) * n, hipMemcpyDeviceToHost));
        double ep = 0, ek = 0;
        for (size_t i = 0; i < n; ++i) {
            ep += e[i];
            const double m = 1.0 / md->massinv_h[t[i]];
            ek += 0.5 * m * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        }
        if (epot) *epot = ep;
        if (ekin) *ekin = ek * MVV2E;
    });
}

}  // extern "C"
