// Training-side kernels of the learned surrogate: the descriptor chain rule that turns per-atom
// energy gradients g = dE_a/dB_a into forces, the adjoint of the force loss with respect to g, and
// the per-configuration energy loss.  All sums are gathers in a fixed order (no atomics), so a
// training step is deterministic.
//
// Layouts (P pairs sorted by centre atom, n atoms, K descriptor components):
//   g, U      [K][n]          dbdr  [K][3][P]  (dB_k(centre)/dr_ij, r_ij = x_j - x_i)
//   s         [3][P]          F, R, Fref  [n][3]
// Pair p = (i = pi[p], j = pj[p]) contributes s_p = sum_k g_ki dbdr_kp to F_i and -s_p to F_j
// (F = -dE/dx with E_i a function of the r_ij of its pairs, as in cpuSnap.cpp:2293-2365).
#include <hip/hip_runtime.h>

#include "sur_launch.hpp"

namespace mdp {

__global__ void k_sur_pair_s(const double* __restrict__ g, const double* __restrict__ dbdr, const int* __restrict__ pi,
                             int K, int n, long P, double* __restrict__ s) {
    const long p = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int i = pi[p];
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int k = 0; k < K; ++k) {
        const double gk = g[static_cast<size_t>(k) * n + i];
        sx += gk * dbdr[(3 * k) * P + p];
        sy += gk * dbdr[(3 * k + 1) * P + p];
        sz += gk * dbdr[(3 * k + 2) * P + p];
    }
    s[p] = sx;
    s[P + p] = sy;
    s[2 * P + p] = sz;
}

__global__ void k_sur_atom_f(const double* __restrict__ s, const int* __restrict__ own_off,
                             const int* __restrict__ in_off, const int* __restrict__ in_idx, long P, int n,
                             const double* __restrict__ fref, double* __restrict__ F, double* __restrict__ R) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int p = own_off[a]; p < own_off[a + 1]; ++p) {
        fx += s[p];
        fy += s[P + p];
        fz += s[2 * P + p];
    }
    for (int q = in_off[a]; q < in_off[a + 1]; ++q) {
        const long p = in_idx[q];
        fx -= s[p];
        fy -= s[P + p];
        fz -= s[2 * P + p];
    }
    if (F) {
        F[3 * a] = fx;
        F[3 * a + 1] = fy;
        F[3 * a + 2] = fz;
    }
    if (R) {
        R[3 * a] = fx - fref[3 * a];
        R[3 * a + 1] = fy - fref[3 * a + 1];
        R[3 * a + 2] = fz - fref[3 * a + 2];
    }
}

// U[k][i] = coef * sum_{p = (i, j)} (R_i - R_j) . dbdr_kp   (= d(coef/2 sum |R|^2)/dg_ki)
__global__ void k_sur_u(const double* __restrict__ R, const double* __restrict__ dbdr, const int* __restrict__ own_off,
                        const int* __restrict__ pj, int K, int n, long P, double coef, double* __restrict__ U) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    constexpr int KMAX = 16;
    double acc[KMAX];
#pragma unroll
    for (int k = 0; k < KMAX; ++k) acc[k] = 0.0;
    const double rx = R[3 * a], ry = R[3 * a + 1], rz = R[3 * a + 2];
    for (int p = own_off[a]; p < own_off[a + 1]; ++p) {
        const int j = pj[p];
        const double dx = rx - R[3 * j], dy = ry - R[3 * j + 1], dz = rz - R[3 * j + 2];
#pragma unroll
        for (int k = 0; k < KMAX; ++k)
            if (k < K) acc[k] += dx * dbdr[(3 * k) * P + p] + dy * dbdr[(3 * k + 1) * P + p] + dz * dbdr[(3 * k + 2) * P + p];
    }
#pragma unroll
    for (int k = 0; k < KMAX; ++k)
        if (k < K) U[static_cast<size_t>(k) * n + a] = coef * acc[k];
}

// per configuration: E_c = sum of its atom energies; alpha_a = coef * 2 (E_c - Eref_c) / N_c^2
// (the derivative of coef * ((E_c - Eref_c) / N_c)^2 with respect to each atom energy)
__global__ void k_sur_energy(const double* __restrict__ e, const int* __restrict__ cfg_off,
                             const double* __restrict__ eref, int C, double coef, double* __restrict__ Ec,
                             double* __restrict__ alpha) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const int a0 = cfg_off[c], a1 = cfg_off[c + 1];
    double E = 0.0;
    for (int a = a0; a < a1; ++a) E += e[a];
    Ec[c] = E;
    if (alpha) {
        const double nc = static_cast<double>(a1 - a0);
        const double al = coef * 2.0 * (E - eref[c]) / (nc * nc);
        for (int a = a0; a < a1; ++a) alpha[a] = al;
    }
}

// per-block partial sums of v[k]^2 (fixed order within a block)
__global__ __launch_bounds__(256) void k_sumsq(const double* __restrict__ v, long n, double* __restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (long k = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; k < n; k += static_cast<long>(gridDim.x) * 256)
        s += v[k] * v[k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (static_cast<int>(threadIdx.x) < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

void launch_sur_pair_s(const double* g, const double* dbdr, const int* pi, int K, int n, long P, double* s,
                       hipStream_t st) {
    if (P > 0) hipLaunchKernelGGL(k_sur_pair_s, dim3((P + 255) / 256), dim3(256), 0, st, g, dbdr, pi, K, n, P, s);
}

void launch_sur_atom_f(const double* s, const int* own_off, const int* in_off, const int* in_idx, long P, int n,
                       const double* fref, double* F, double* R, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(k_sur_atom_f, dim3((n + 255) / 256), dim3(256), 0, st, s, own_off, in_off, in_idx, P, n, fref,
                           F, R);
}

void launch_sur_u(const double* R, const double* dbdr, const int* own_off, const int* pj, int K, int n, long P,
                  double coef, double* U, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(k_sur_u, dim3((n + 255) / 256), dim3(256), 0, st, R, dbdr, own_off, pj, K, n, P, coef, U);
}

void launch_sur_energy(const double* e, const int* cfg_off, const double* eref, int C, double coef, double* Ec,
                       double* alpha, hipStream_t st) {
    if (C > 0)
        hipLaunchKernelGGL(k_sur_energy, dim3((C + 127) / 128), dim3(128), 0, st, e, cfg_off, eref, C, coef, Ec, alpha);
}

void launch_sumsq(const double* v, long n, double* part, int nblocks, hipStream_t st) {
    hipLaunchKernelGGL(k_sumsq, dim3(nblocks), dim3(256), 0, st, v, n, part);
}

}  // namespace mdp
